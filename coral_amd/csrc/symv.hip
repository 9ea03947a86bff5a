// Level-2 SYMV / HEMV and TRMV for gfx950: y := alpha*A*x + beta*y with A symmetric
// (s/d) or Hermitian (c/z), and x := op(A)*x with A triangular, each read from the
// `uplo` triangle only.
//
// CORAL walks the stored triangle once per column (lda == n path,
// coral-aarch64/src/level2/chemv.rs:142-231): every stored a_ij feeds y_i (axpy)
// and y_j (conjugated dot), the diagonal contributes Re(a_jj) only (:178-182), and
// the other triangle is never used.  TRMV (coral-aarch64/src/level2/dtrmv.rs:58,
// dtrlmv.rs:47-469, dtrumv.rs:48-460, ctrmv.rs, ztrmv.rs) keeps one of those two
// products.  Here every stored element is streamed from HBM exactly once, by
// wavefronts that each own a row group and a run of its stored columns (the
// streaming kernel below), and a second kernel adds the partial sums in a fixed
// order (beta == 0 overwrites y, as chemv.rs:135-139 does).
#include "coral_common.hpp"

#include <algorithm>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <stdint.h>

#include <type_traits>

namespace coral {
namespace sym {

typedef uint32_t sv4u __attribute__((ext_vector_type(4)));

// ---- by 64 x 64 blocks (orders below kStreamMinN) ----
// The stored triangle is cut into NB x NB blocks, one workgroup per block, and each
// block is streamed from HBM exactly once (a diagonal block is read whole; its
// unstored half is masked out):
//   off-diagonal block (r, c):  u = A_rc * x_c  -> partial for row block r
//                               v = A_rc^H * x_r -> partial for row block c
//   diagonal block (b, b):      the stored half of the block, both roles.
// Every row block receives exactly nb partials (one per partner block), stored in
// Q[r][c][NB]; a second kernel adds them in c order and applies alpha/beta.  TRMV
// keeps one of the two products: op = N takes the block's rows (partial of row
// block r from partner c), op = T / C its columns (partial of row block c from
// partner r); a row block's partners are then the blocks on one side of the
// diagonal.
template <typename E> __device__ __forceinline__ E real_part(E a) {
  if constexpr (elem_traits<E>::cplx) { a.y = 0; return a; }
  else return a;
}

constexpr int NB = 64;  // block edge

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

// ---- TRMV by blocks ----
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


// ---- streaming SYMV / HEMV ----
// The rows are cut into row groups of RW = 64 * NV * EV rows (EV elements per
// 16-byte vector).  A wavefront streams one row group across a run of columns as
// GEMV N does: lane l holds rows rbase + v*64*EV + l*EV + e (v < NV, e < EV), reads
// NV 16-byte vectors of each column, and keeps the next four columns' loads in
// flight while it uses the current four.  Per stored element a_ij:
//   u_i += a_ij x_j        kept in the lane across columns (row sums);
//   v_j += a_ij' x_i       (a' = conj(a) for Hermitian) summed over the lane's rows,
//                          then over the wavefront's 64 lanes (reduce4), which
//                          leaves column 4c + q of a 64-column window in lane 16q + c.
// A row group's stored columns (lower: 0 .. min(n, group end) - 1; upper: group
// start .. n - 1) are concatenated over the row groups, group-major, and that
// sequence is cut into equal runs of W columns, one per wavefront, so every
// wavefront streams the same share of the triangle.  A run that crosses a group
// boundary is processed as one piece per group; piece (run r, group g) has index
// r + g (pieces are ordered and each step to the next one advances r, g or both) and
// writes its row sums to U[r + g][RW].  Row group g writes its column sums to V[g][j]
// for every column it covers: lower, j < (g + 1) * RW; upper, j >= g * RW.  A second
// kernel adds, for row i of group g, the U pieces of group g in order and the V
// rows of the groups that hold column i's stored elements in order, and applies
// alpha/beta.
// Orders from which the streaming kernel replaces the block kernel, for elements of
// up to 8 bytes; complex double always goes by blocks.  Measured on one box, whole
// call, blocks / stream: n = 16384 SSYMV 94.8 / 92.4 us, DSYMV 192.4 / 184.1, CHEMV
// 191.0 / 184.5, DTRMV N 180.8 / 174.8, ZHEMV 375.5 / 378.6; n = 12288 even (within
// 3%); n = 8192 and 10240 the blocks win by 3-15% (a wavefront's run is short enough
// that its first loads' latency and the second launch weigh).
constexpr int64_t kStreamMinN = 12288;

template <typename E, int NV> struct SymvShape {
  static constexpr int EV = 16 / sizeof(E);
  static constexpr int RW = 64 * NV * EV;  // rows of a row group (one wavefront)
};

struct SymvGeom {
  int64_t n, rw, ng;  // matrix order, rows per group, groups
  int64_t tc, w;      // concatenated stored columns of all groups, columns per run
  int upper;
};

// First concatenated column of row group g (0 <= g <= ng): lower groups before the
// last are full, group t holding (t + 1) rw columns; upper group t holds n - t rw.
__host__ __device__ inline int64_t group_first(const SymvGeom &g, int64_t t) {
  if (t >= g.ng) return g.tc;
  return g.upper ? t * g.n - g.rw * t * (t - 1) / 2 : g.rw * t * (t + 1) / 2;
}

// The row group whose concatenated columns hold column a (0 <= a < tc).
__device__ __forceinline__ int64_t group_of(const SymvGeom &g, int64_t a) {
  double est;
  if (g.upper) {
    const double p = (double)g.n + 0.5 * (double)g.rw;
    const double d = p * p - 2.0 * (double)g.rw * (double)a;
    est = (p - sqrt(d > 0.0 ? d : 0.0)) / (double)g.rw;
  } else {
    est = (sqrt(1.0 + 8.0 * (double)a / (double)g.rw) - 1.0) * 0.5;
  }
  int64_t t = (int64_t)est;
  t = t < 0 ? 0 : (t >= g.ng ? g.ng - 1 : t);
  while (t > 0 && group_first(g, t) > a) --t;
  while (t + 1 < g.ng && group_first(g, t + 1) <= a) ++t;
  return t;
}

// Cross-lane moves on the 32-bit words of an element (an element of 8 or 16 bytes
// moves word by word; the words of one lane stay together).
template <typename E> __device__ __forceinline__ void swap_halves32(E &a, E &b) {
  // v_permlane32_swap: lanes 32..63 of a trade places with lanes 0..31 of b
  constexpr int NW = sizeof(E) / 4;
  uint32_t wa[NW], wb[NW];
  __builtin_memcpy(wa, &a, sizeof(E));
  __builtin_memcpy(wb, &b, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const auto r = __builtin_amdgcn_permlane32_swap(wa[i], wb[i], false, false);
    wa[i] = r[0];
    wb[i] = r[1];
  }
  __builtin_memcpy(&a, wa, sizeof(E));
  __builtin_memcpy(&b, wb, sizeof(E));
}
template <typename E> __device__ __forceinline__ void swap_rows16(E &a, E &b) {
  // v_permlane16_swap: the odd 16-lane rows of a trade places with the even rows of b
  constexpr int NW = sizeof(E) / 4;
  uint32_t wa[NW], wb[NW];
  __builtin_memcpy(wa, &a, sizeof(E));
  __builtin_memcpy(wb, &b, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) {
    const auto r = __builtin_amdgcn_permlane16_swap(wa[i], wb[i], false, false);
    wa[i] = r[0];
    wb[i] = r[1];
  }
  __builtin_memcpy(&a, wa, sizeof(E));
  __builtin_memcpy(&b, wb, sizeof(E));
}
// DPP lane permutation within 16-lane rows: 0xB1 = quad_perm [1,0,3,2] (lane ^ 1),
// 0x4E = quad_perm [2,3,0,1] (lane ^ 2), 0x141 = row_half_mirror (i <-> 7 - i in each
// 8 lanes), 0x140 = row_mirror (i <-> 15 - i).
template <int CTRL, typename E> __device__ __forceinline__ E dpp_move(E v) {
  constexpr int NW = sizeof(E) / 4;
  uint32_t w[NW];
  __builtin_memcpy(w, &v, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[i], CTRL, 0xF, 0xF, false);
  __builtin_memcpy(&v, w, sizeof(E));
  return v;
}

// Sums four per-lane column partials over the wavefront: afterwards every lane of
// 16-lane row q holds the full sum of column q.  The two 32-lane halves trade the
// columns they do not keep (lanes 0..31 keep 0 and 1, lanes 32..63 keep 2 and 3), the
// rows then trade again (even rows keep the first of their pair, odd rows the second)
// and each row is summed with four DPP steps, each of which pairs lanes whose groups
// hold equal partial sums (quads, halves of 8, halves of 16).
template <typename E> __device__ __forceinline__ E reduce4(E c0, E c1, E c2, E c3) {
  swap_halves32(c0, c2);
  swap_halves32(c1, c3);
  E k0 = eadd(c0, c2), k1 = eadd(c1, c3);  // lanes < 32: columns 0, 1; lanes >= 32: 2, 3
  swap_rows16(k0, k1);
  E m = eadd(k0, k1);                      // row q: column q
  m = eadd(m, dpp_move<0xB1>(m));
  m = eadd(m, dpp_move<0x4E>(m));
  m = eadd(m, dpp_move<0x141>(m));
  m = eadd(m, dpp_move<0x140>(m));
  return m;
}

// An element with each 32-bit word ANDed with a mask (all ones keeps it, zero clears
// it); for complex elements the real and imaginary words take their own masks.
template <typename E> __device__ __forceinline__ E and_mask(E v, uint32_t mre, uint32_t mim) {
  constexpr int NW = sizeof(E) / 4;
  uint32_t w[NW];
  __builtin_memcpy(w, &v, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) w[i] &= (elem_traits<E>::cplx && i >= NW / 2) ? mim : mre;
  __builtin_memcpy(&v, w, sizeof(E));
  return v;
}
// Each 32-bit word from b where m is all ones, from a where it is zero.
template <typename E> __device__ __forceinline__ E select_bits(E a, E b, uint32_t m) {
  constexpr int NW = sizeof(E) / 4;
  uint32_t wa[NW], wb[NW];
  __builtin_memcpy(wa, &a, sizeof(E));
  __builtin_memcpy(wb, &b, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) wa[i] = (wa[i] & ~m) | (wb[i] & m);
  __builtin_memcpy(&a, wa, sizeof(E));
  return a;
}

// What a streamed row group computes from its stored elements a_ij (d = i - j for
// the upper triangle, j - i for the lower; stored means d <= 0):
//   SM_SYM  u_i += a_ij x_j (d <= 0; a Hermitian diagonal with its real part only)
//           and v_j += a_ij' x_i (d < 0): SYMV / HEMV;
//   SM_ROWS u_i += a_ij x_j (d <= 0; a unit diagonal reads as 1): TRMV op = N;
//   SM_COLS v_j += a_ij' x_i (d <= 0; unit diagonal as 1): TRMV op = T / C.
// a' = conj(a) when CONJ (HEMV, TRMV C), a otherwise.
enum StreamMode { SM_SYM = 0, SM_ROWS = 1, SM_COLS = 2 };

// One wavefront's piece: columns [ja, jb) of the row group at rbase.  FULL: every
// row of the group lies inside the matrix and columns are 16-byte aligned, so each
// column vector is a buffer load from a scalar descriptor at the segment's first row
// plus the lane's fixed 32-bit offset (no vector register holds an address: a 64-bit
// address register that a load in flight also writes makes the next group wait for
// that load).  Two register sets alternate: the next four columns' loads (from
// column indices clamped to jb - 1, so no load sits behind a branch) are issued
// before the current four are used.  Column and row indices are 32-bit (an order of
// 2^31 would need 2^62 elements), so the loop's compares are scalar.
template <typename E, int MODE, bool CONJ, int NV, bool FULL>
__device__ __forceinline__ void stream_rows(const E *__restrict__ A, int64_t lda, const E *__restrict__ x,
                                            int64_t incx, int n, bool upper, bool unit, int rbase, int ja, int jb,
                                            const E (&xr)[NV][16 / sizeof(E)], E (&acc)[NV][16 / sizeof(E)],
                                            E *__restrict__ vout, int lane) {
  constexpr int EV = 16 / sizeof(E), RW = 64 * NV * EV;
  constexpr bool DO_U = MODE != SM_COLS, DO_V = MODE != SM_ROWS;
  int row[NV];
  uint32_t boff[NV];  // the lane's byte offset from the row group's first row
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    row[v] = rbase + v * 64 * EV + lane * EV;
    boff[v] = (uint32_t)((v * 64 * EV + lane * EV) * (int)sizeof(E));
  }
  typedef const __attribute__((address_space(4))) uint32_t *xcptr;  // x_j by scalar loads
  auto load = [&](int j0, E (&av)[4][NV][EV], E (&xj)[4]) {
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + c < jb ? j0 + c : jb - 1;
      if constexpr (DO_U) {
        uint32_t wv[sizeof(E) / 4];
        const xcptr xp = (xcptr)(uintptr_t)(x + (int64_t)j * incx);
#pragma unroll
        for (int i = 0; i < (int)(sizeof(E) / 4); ++i) wv[i] = xp[i];
        __builtin_memcpy(&xj[c], wv, sizeof(E));
      } else {
        xj[c] = ezero<E>();
      }
      const E *col = A + (int64_t)j * lda;
      if constexpr (FULL) {
        const __amdgpu_buffer_rsrc_t rs =
            __builtin_amdgcn_make_buffer_rsrc((void *)(col + rbase), (short)0, RW * (int)sizeof(E), 0x00020000);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          const sv4u pk = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)boff[v], 0, 2 /* nontemporal */);
          __builtin_memcpy(&av[c][v][0], &pk, 16);
        }
      } else {
#pragma unroll
        for (int v = 0; v < NV; ++v)
#pragma unroll
          for (int e = 0; e < EV; ++e) av[c][v][e] = row[v] + e < n ? col[row[v] + e] : ezero<E>();
      }
    }
  };
  const E one = make_e<E>(1.0, 0.0);
  E vsum = ezero<E>();
  // MASK: the group meets the diagonal block (some element is not stored, or lies on
  // the diagonal) or runs past jb.  Its element masks are arithmetic (an arithmetic
  // shift of a difference of indices), so the group has neither branches nor
  // compares.  A column at or past jb (a repeat of column jb - 1) is cleared from u by
  // a uniform mask (its column sum is never stored).  Without MASK every element is a
  // plain stored off-diagonal element.
  auto group = [&](auto mask_tag, int j0, const E (&av)[4][NV][EV], const E (&xj)[4]) {
    constexpr bool MASK = decltype(mask_tag)::value;
    E vp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int j = j0 + c;
      const uint32_t live = j < jb ? 0xFFFFFFFFu : 0u;  // uniform
      E t = ezero<E>();
#pragma unroll
      for (int v = 0; v < NV; ++v)
#pragma unroll
        for (int e = 0; e < EV; ++e) {
          const E el = av[c][v][e];
          if constexpr (!MASK) {
            if constexpr (DO_U) acc[v][e] = emad(acc[v][e], el, xj[c]);
            if constexpr (DO_V) t = emad<CONJ>(t, el, xr[v][e]);
          } else {
            const int i = row[v] + e;
            const int d = upper ? i - j : j - i;
            const uint32_t m_le = (uint32_t)((d - 1) >> 31);  // d <= 0: stored
            const uint32_t m_lt = (uint32_t)(d >> 31);        // d < 0: stored, off the diagonal
            const uint32_t m_eq = m_le & ~m_lt;               // the diagonal
            if constexpr (MODE == SM_SYM) {
              const E eu = and_mask(el, m_le & live, (CONJ ? m_lt : m_le) & live);
              const E ev = and_mask(el, m_lt, m_lt);
              acc[v][e] = emad(acc[v][e], eu, xj[c]);
              t = emad<CONJ>(t, ev, xr[v][e]);
            } else {
              E ed = and_mask(el, m_le, m_le);
              if (unit) ed = select_bits(ed, one, m_eq);
              if constexpr (MODE == SM_ROWS) acc[v][e] = emad(acc[v][e], and_mask(ed, live, live), xj[c]);
              else t = emad<CONJ>(t, ed, xr[v][e]);
            }
          }
        }
      vp[c] = t;
    }
    if constexpr (DO_V) return reduce4(vp[0], vp[1], vp[2], vp[3]);  // lane 16q + r: column j0 + q
    else return ezero<E>();
  };
  auto use = [&](int j0, const E (&av)[4][NV][EV], const E (&xj)[4]) {
    const int cg = ((j0 - ja) >> 2) & 15;
    const bool masked = (upper ? j0 < rbase + RW : j0 + 4 > rbase) || j0 + 4 > jb;  // uniform
    const E m = masked ? group(std::true_type{}, j0, av, xj) : group(std::false_type{}, j0, av, xj);
    if constexpr (DO_V) {
      if ((lane & 15) == cg) vsum = m;
      if (cg == 15 || j0 + 4 >= jb) {  // a 64-column window is complete: store it
        const int col = j0 - 4 * cg + 4 * (lane & 15) + (lane >> 4);
        if (col < jb) vout[col] = vsum;
        vsum = ezero<E>();
      }
    }
  };
  if (ja >= jb) return;
  E a0[4][NV][EV], a1[4][NV][EV], x0[4], x1[4];
  load(ja, a0, x0);
  for (int j0 = ja;;) {
    load(j0 + 4, a1, x1);
    use(j0, a0, x0);
    j0 += 4;
    if (j0 >= jb) break;
    load(j0 + 4, a0, x0);
    use(j0, a1, x1);
    j0 += 4;
    if (j0 >= jb) break;
  }
}

// One run of W concatenated columns per wavefront (four per workgroup, which share
// nothing: no LDS, no barrier).
template <typename E, int MODE, bool CONJ, int NV>
__global__ __launch_bounds__(256) void tri_stream_kernel(SymvGeom geo, const E *__restrict__ A, int64_t lda,
                                                         const E *__restrict__ x, int64_t incx, E *__restrict__ U,
                                                         E *__restrict__ V, int vec_ok, int unit) {
  using SH = SymvShape<E, NV>;
  constexpr int EV = SH::EV, RW = SH::RW;
  const int lane = threadIdx.x & 63;
  // the run index as a scalar: everything derived from it (the row group, its column
  // range, the mask decisions) is then uniform, and the column loop a scalar loop
  const int64_t run = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t n = geo.n;
  const bool upper = geo.upper;
  int64_t a = run * geo.w;
  if (a >= geo.tc) return;
  const int64_t b = a + geo.w < geo.tc ? a + geo.w : geo.tc;
  for (int64_t g = group_of(geo, a); a < b; ++g) {
    const int64_t first = group_first(geo, g), end = group_first(geo, g + 1);
    const int64_t pe = b < end ? b : end;
    const int64_t rbase = g * RW;
    const int64_t ja = (upper ? rbase : 0) + (a - first), jb = (upper ? rbase : 0) + (pe - first);
    a = pe;
    E xr[NV][EV], acc[NV][EV];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      const int64_t r = rbase + v * 64 * EV + lane * EV;
#pragma unroll
      for (int e = 0; e < EV; ++e) {
        xr[v][e] = (MODE != SM_ROWS && r + e < n) ? x[(r + e) * incx] : ezero<E>();
        acc[v][e] = ezero<E>();
      }
    }
    E *vout = V + g * n;
    if (vec_ok && rbase + RW <= n)
      stream_rows<E, MODE, CONJ, NV, true>(A, lda, x, incx, (int)n, upper, unit, (int)rbase, (int)ja, (int)jb, xr,
                                           acc, vout, lane);
    else
      stream_rows<E, MODE, CONJ, NV, false>(A, lda, x, incx, (int)n, upper, unit, (int)rbase, (int)ja, (int)jb, xr,
                                            acc, vout, lane);
    if constexpr (MODE != SM_COLS) {
      E *uout = U + (run + g) * RW;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const int64_t r = rbase + v * 64 * EV + lane * EV;
#pragma unroll
        for (int e = 0; e < EV; ++e)
          if (r + e < n) uout[v * 64 * EV + lane * EV + e] = acc[v][e];
      }
    }
  }
}

// out[i] := the row sums of the pieces of i's row group (modes with row sums) plus
// the column sums of the row groups that hold stored elements of column i (modes
// with column sums), then alpha * sum + beta * out[i] (SM_SYM) or the sum itself
// (TRMV).  A workgroup takes 64 rows (one row group); its four wavefronts sum
// consecutive quarters of the pieces and of the row groups in order, eight loads in
// flight, and the quarter sums are added in order, so the result does not depend on
// scheduling.
template <typename E, int MODE, int NV>
__global__ __launch_bounds__(256) void tri_stream_reduce_kernel(SymvGeom geo, const E *__restrict__ U,
                                                                const E *__restrict__ V, E alpha, E beta,
                                                                int beta_zero, E *out, int64_t inc) {
  constexpr int RW = SymvShape<E, NV>::RW;
  __shared__ E part[4][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  const int64_t i0 = (int64_t)blockIdx.x * 64, i = i0 + lane;
  const int64_t g = i0 / RW;  // uniform: RW is a multiple of 64
  E acc = ezero<E>();
  if (i < geo.n) {
    if constexpr (MODE != SM_COLS) {
      const int64_t p0 = group_first(geo, g) / geo.w + g, p1 = (group_first(geo, g + 1) - 1) / geo.w + g + 1;
      const int64_t pq = (p1 - p0 + 3) / 4, pa = p0 + q * pq, pb = pa + pq < p1 ? pa + pq : p1;
      const E *ub = U + (i - g * RW);
      int64_t p = pa;
      for (; p + 8 <= pb; p += 8) {
        E v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = ub[(p + k) * RW];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = eadd(acc, v[k]);
      }
      for (; p < pb; ++p) acc = eadd(acc, ub[p * RW]);
    }
    if constexpr (MODE != SM_ROWS) {
      const int64_t g0 = geo.upper ? 0 : g, g1 = geo.upper ? g + 1 : geo.ng;
      const int64_t gq = (g1 - g0 + 3) / 4, ga = g0 + q * gq, gb = ga + gq < g1 ? ga + gq : g1;
      const E *vb = V + i;
      int64_t t = ga;
      for (; t + 8 <= gb; t += 8) {
        E v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) v[k] = vb[(t + k) * geo.n];
#pragma unroll
        for (int k = 0; k < 8; ++k) acc = eadd(acc, v[k]);
      }
      for (; t < gb; ++t) acc = eadd(acc, vb[t * geo.n]);
    }
  }
  part[q][lane] = acc;
  __syncthreads();
  if (q == 0 && i < geo.n) {
    const E t = eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane]));
    E *op = out + i * inc;
    if constexpr (MODE == SM_SYM) *op = axpby(alpha, t, beta, op, beta_zero);
    else *op = t;
  }
}

template <typename E>
__global__ void symv_beta_kernel(int64_t n, E beta, int beta_zero, E *y, int64_t incy) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    E *yp = y + i * incy;
    *yp = beta_zero ? ezero<E>() : emul(beta, *yp);
  }
}

// Launches the streaming product and its reduce: about wg_per_cu workgroups of four
// wavefronts per CU, one run of W >= 16 columns per wavefront.  out / inc: y for
// SYMV / HEMV (alpha, beta applied), x for TRMV (overwritten once every run has read
// it: the reduce is the second launch).
template <typename E, int MODE, bool CONJ, int NV>
static int tri_stream(int32_t uplo, int64_t n, const E *A, int64_t lda, const E *x, int64_t incx, E alpha, E beta,
                      E *out, int64_t inc, int unit, int vec_ok, int wg_per_cu, hipStream_t s) {
  using SH = SymvShape<E, NV>;
  SymvGeom geo;
  geo.n = n;
  geo.rw = SH::RW;
  geo.ng = cdiv(n, SH::RW);
  geo.upper = uplo == CORAL_UPPER_TRIANGULAR;
  geo.tc = geo.upper ? geo.ng * n - geo.rw * geo.ng * (geo.ng - 1) / 2 : geo.rw * (geo.ng - 1) * geo.ng / 2 + n;
  geo.w = std::max<int64_t>(16, cdiv(geo.tc, 4 * (int64_t)wg_per_cu * num_cus()));
  const int64_t runs = cdiv(geo.tc, geo.w);
  const int64_t pieces = runs + geo.ng - 1;
  const size_t u_bytes = MODE == SM_COLS ? 0 : (sizeof(E) * (size_t)(pieces * SH::RW) + 255) & ~size_t(255);
  const size_t v_bytes = MODE == SM_ROWS ? 0 : sizeof(E) * (size_t)(geo.ng * n);
  void *ws = nullptr;
  CORAL_TRY(workspace(s, WS_AUX2, u_bytes + v_bytes + 256, &ws));
  E *U = (E *)ws, *V = (E *)((char *)ws + u_bytes);
  hipLaunchKernelGGL((tri_stream_kernel<E, MODE, CONJ, NV>), dim3((unsigned)cdiv(runs, 4)), dim3(256), 0, s, geo, A,
                     lda, x, incx, U, V, vec_ok, unit);
  CORAL_LAUNCH_CHECK();
  hipLaunchKernelGGL((tri_stream_reduce_kernel<E, MODE, NV>), dim3((unsigned)cdiv(n, 64)), dim3(256), 0, s, geo,
                     (const E *)U, (const E *)V, alpha, beta, (int)eiszero(beta), out, inc);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// The shape of the streaming kernels: vectors per column (NV) and workgroups per CU.
// Measured on 16384 SSYMV / DSYMV / CHEMV and 8192 ZHEMV, whole call: two vectors and
// one workgroup per CU (a run of tc / (4 x CUs) columns per wavefront) is fastest;
// more workgroups shorten the runs, four vectors cost registers.
// CORAL_SYMV_CFG=<vectors>,<workgroups per CU> overrides it.
// The order from which SYMV / HEMV / TRMV on elements E stream (kStreamMinN, or never
// for complex double); CORAL_SYMV_STREAM_MIN_N overrides it (tests run the streaming
// kernel at small orders with it).
template <typename E> static int64_t stream_min_n() {
  if (const char *e = getenv("CORAL_SYMV_STREAM_MIN_N")) return atoll(e);
  return sizeof(E) <= 8 ? kStreamMinN : INT64_MAX;
}

static void stream_cfg(int &nv, int &wg) {
  nv = 2;
  wg = 1;
  if (const char *e = getenv("CORAL_SYMV_CFG")) sscanf(e, "%d,%d", &nv, &wg);
  if (nv != 4) nv = 2;
  if (wg < 1 || wg > 16) wg = 1;
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
  const int vec_ok = ((uintptr_t)A % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  if (n >= stream_min_n<E>()) {
    int nv, wg;
    stream_cfg(nv, wg);
    constexpr int M = SM_SYM;
    return nv == 4 ? tri_stream<E, M, HERM, 4>(uplo, n, A, lda, x, incx, alpha, beta, y, incy, 0, vec_ok, wg, s)
                   : tri_stream<E, M, HERM, 2>(uplo, n, A, lda, x, incx, alpha, beta, y, incy, 0, vec_ok, wg, s);
  }
  const int64_t nb = cdiv(n, NB);
  void *Q = nullptr;
  CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)(nb * nb * NB), &Q));
  const int64_t nblocks = nb * (nb + 1) / 2;
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
// x := op(A) x: from kStreamMinN on, the stored triangle is streamed like SYMV's and
// one of the two products kept: op = N the row sums (SM_ROWS), op = T / C the column
// sums (SM_COLS, conjugated for C); the diagonal counts in it, as 1 for a unit
// diagonal, whose stored value is never used.  Smaller orders go by 64 x 64 blocks.
// Either way x is overwritten by the reduce, after every block or run has read it.
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
  // real routines map ConjugateTranspose to Transpose (coral-aarch64/src/level2/dtrmv.rs:58-79)
  const int op = (trans == CORAL_CONJUGATE_TRANSPOSE && !elem_traits<E>::cplx) ? 1 : trans;
  const int unit = diag == CORAL_UNIT_DIAGONAL;
  const int vec_ok = ((uintptr_t)a % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  if (n >= stream_min_n<E>()) {
    int nv, wg;
    stream_cfg(nv, wg);
    const E one = make_e<E>(1.0, 0.0), zero = ezero<E>();
#define CORAL_TRMV_CASE(MODE, CONJ)                                                                                  \
    return nv == 4 ? tri_stream<E, MODE, CONJ, 4>(uplo, n, a, lda, x, incx, one, zero, x, incx, unit, vec_ok, wg, s) \
                   : tri_stream<E, MODE, CONJ, 2>(uplo, n, a, lda, x, incx, one, zero, x, incx, unit, vec_ok, wg, s)
    if (op == 0) CORAL_TRMV_CASE(SM_ROWS, false);
    if (op == 1) CORAL_TRMV_CASE(SM_COLS, false);
    CORAL_TRMV_CASE(SM_COLS, true);
#undef CORAL_TRMV_CASE
  }
  const int upper = uplo == CORAL_UPPER_TRIANGULAR;
  const int64_t nb = cdiv(n, NB);
  void *Q = nullptr;
  CORAL_TRY(workspace(s, WS_AUX2, sizeof(E) * (size_t)(nb * nb * NB), &Q));
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
