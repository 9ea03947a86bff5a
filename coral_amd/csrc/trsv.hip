// Level-2 TRSV for gfx950, plus batched DTRSV.
//
// CORAL solves op(A) x = b in NB = 8 diagonal blocks: a scalar substitution on the
// diagonal block (coral-aarch64/src/level2/dtrlsv.rs:65-139) followed by a fused
// `daxpyf` update of the rows below (:230-257, :316-347); transposed and upper
// variants walk the blocks in the other direction (dtrlsv.rs:356-413, dtrusv.rs).
// Complex ConjTranspose conjugates in the substitution and divides by conj(a_ii)
// with den = re^2 + im^2 (ctrlsv.rs:189, 205-214).
//
// Here every variant is mapped onto one forward (lower) solve of a virtual matrix
// M: index i of M is row i of op(A) for forward solves and row n-1-i for backward
// solves, so M(i, c) (c <= i) reads A[.] through the op and the index reversal.
// The solve is one persistent launch:
//   * workgroups take row blocks of NB = 64 rows from an atomic ticket, so a
//     workgroup only ever waits on blocks owned by workgroups already running;
//   * a row block accumulates M(block, j) x_j over earlier blocks j as soon as
//     block j's values are published (tile loads are issued before the wait);
//   * the diagonal block's inverse is formed while the first tiles load (it
//     depends on nothing), the block's right-hand side minus the off-diagonal sums
//     is multiplied by it, in the element precision (FP64 for D/Z), and x_b is
//     published through a sentinel-filled staging vector whose words are their own
//     ready flags (see Stage below);
//   * from the seventh block on, the last dependency (ND = 1, the default for
//     elements up to 8 bytes) or two (ND = 2) are precomputed as products with
//     that inverse, so the step between a predecessor's publish and this block's
//     is one 64-long dot product per row in one wavefront.
#include "coral_common.hpp"

#include <algorithm>
#include <stdlib.h>

namespace coral {
namespace tri {

constexpr int NB = 64;
// Blocks below this index solve their last dependency directly (trsv_kernel).
constexpr int kDirect = 6;

template <typename E>
struct TriView {
  const E *A;
  int64_t lda, n;
  int trans;  // 0 N, 1 T, 2 C
  int fwd;    // 1: forward (effective lower), 0: backward
  __device__ __forceinline__ int64_t map(int64_t i) const { return fwd ? i : n - 1 - i; }
  __device__ __forceinline__ E get(int64_t i, int64_t c) const {
    const int64_t ri = map(i), rc = map(c);
    if (trans == 0) return A[ri + rc * lda];
    E v = A[rc + ri * lda];
    return trans == 2 ? econj(v) : v;
  }
  // The stored element behind M(i, c) when the layout is known at compile time
  // (RC: trans == 0).  For trans != 0 this is A's element before conjugation: a
  // select on the value right after its load would make every prefetched tile
  // load wait for completion, so ConjTrans conjugates later, in store_tile.
  template <bool RC>
  __device__ __forceinline__ E get_raw(int64_t i, int64_t c) const {
    const int64_t ri = map(i), rc = map(c);
    if constexpr (RC) return A[ri + rc * lda];
    else return A[rc + ri * lda];
  }
  // true when consecutive i are adjacent in memory (coalesced thread-per-row loads)
  __device__ __forceinline__ bool row_contig() const { return trans == 0; }
};

// Loads tile (rows i0.., cols c0..) of M into registers, 16 elements per thread.
// diag: only the strictly lower part (c < i) counts; the rest is zero.
// RCT: -1 decides the layout at run time; 0 / 1 fix it at compile time (1: trans == 0).
//
// With RCT >= 0 every load is unconditional, from an index clamped into the n x n
// matrix, and the registers hold A's elements as stored; store_tile applies the
// mask (rows/columns past n, the upper part of a diagonal tile) and ConjTrans's
// conjugate when it writes the tile to LDS, so the two are always used as a pair.
// Any use of a loaded value here (a select, a conjugate, a guarded load merged
// into another register) makes the compiler wait for that load at once, which
// would put a full memory round trip on every prefetched tile.
template <int RCT = -1, typename E>
__device__ __forceinline__ void load_tile(const TriView<E> &v, int64_t i0, int64_t c0, bool diag, E (&reg)[16]) {
  const int t = threadIdx.x;
  const bool rc = RCT < 0 ? v.row_contig() : (RCT == 1);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    // row-contiguous: lanes walk rows; otherwise lanes walk columns
    const int r = rc ? (t & 63) : ((t >> 6) + 4 * k);
    const int c = rc ? ((t >> 6) + 4 * k) : (t & 63);
    const int64_t i = i0 + r, cc = c0 + c;
    if constexpr (RCT < 0) {
      const bool ok = i < v.n && cc < v.n && (!diag || c < r);
      reg[k] = ok ? v.get(i, cc) : ezero<E>();
    } else {
      reg[k] = v.template get_raw<RCT == 1>(i < v.n ? i : v.n - 1, cc < v.n ? cc : v.n - 1);
    }
  }
}

template <int RCT = -1, typename E>
__device__ __forceinline__ void store_tile(const TriView<E> &v, E (*T)[NB + 1], const E (&reg)[16], int64_t i0,
                                           int64_t c0, bool diag) {
  const int t = threadIdx.x;
  const bool rc = RCT < 0 ? v.row_contig() : (RCT == 1);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const int r = rc ? (t & 63) : ((t >> 6) + 4 * k);
    const int c = rc ? ((t >> 6) + 4 * k) : (t & 63);
    // column-major tile: T[c][r] = M(i0 + r, c0 + c)
    if constexpr (RCT < 0) {
      T[c][r] = reg[k];
    } else {
      const bool ok = i0 + r < v.n && c0 + c < v.n && (!diag || c < r);
      E e = reg[k];
      if constexpr (RCT == 0) e = v.trans == 2 ? econj(e) : e;
      T[c][r] = ok ? e : ezero<E>();
    }
  }
}

template <typename E>
__device__ __forceinline__ E ediv(E a, E d) {
  if constexpr (elem_traits<E>::cplx) {
    // (a.x + i a.y) / (d.x + i d.y) = a * conj(d) / (d.x^2 + d.y^2)   (ctrlsv.rs:205-214)
    auto den = d.x * d.x + d.y * d.y;
    E r;
    r.x = (a.x * d.x + a.y * d.y) / den;
    r.y = (a.y * d.x - a.x * d.y) / den;
    return r;
  } else {
    return a / d;
  }
}

template <typename E>
__device__ __forceinline__ E eshfl(E v, int src) {
  if constexpr (elem_traits<E>::cplx) { v.x = __shfl(v.x, src, 64); v.y = __shfl(v.y, src, 64); return v; }
  else return __shfl(v, src, 64);
}
template <typename E> __device__ __forceinline__ E eneg(E a) {
  if constexpr (elem_traits<E>::cplx) { a.x = -a.x; a.y = -a.y; return a; } else return -a;
}
template <typename E> __device__ __forceinline__ E esub(E a, E b) { return eadd(a, eneg(b)); }
template <typename E> __device__ __forceinline__ E eone() {
  if constexpr (elem_traits<E>::cplx) { E a{}; a.x = 1; return a; } else return E(1);
}

// Broadcast lane `l` (wavefront-uniform) through the scalar unit.
__device__ __forceinline__ float rlane(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double rlane(double v, int l) {
  const uint64_t u = (uint64_t)__double_as_longlong(v);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ __forceinline__ float2 rlane(float2 v, int l) { return make_float2(rlane(v.x, l), rlane(v.y, l)); }
__device__ __forceinline__ double2 rlane(double2 v, int l) { return make_double2(rlane(v.x, l), rlane(v.y, l)); }

// a / d from a precomputed r = 1/d: q = a*r corrected by one Newton step
// (q + (a - q*d)*r), which gives the rounded quotient for real types.
__device__ __forceinline__ float qdiv(float a, float d, float r) { float q = a * r; return fmaf(fmaf(-q, d, a), r, q); }
__device__ __forceinline__ double qdiv(double a, double d, double r) { double q = a * r; return fma(fma(-q, d, a), r, q); }
__device__ __forceinline__ float2 qdiv(float2 a, float2 d, float2 r) { (void)d; return emul(a, r); }
__device__ __forceinline__ double2 qdiv(double2 a, double2 d, double2 r) { (void)d; return emul(a, r); }
template <typename E> __device__ __forceinline__ E erecip(E d) {
  if constexpr (elem_traits<E>::cplx) {
    auto den = d.x * d.x + d.y * d.y;  // 1/d = conj(d) / (re^2 + im^2)  (ctrlsv.rs:205-214)
    E r;
    r.x = d.x / den;
    r.y = -d.y / den;
    return r;
  } else {
    return E(1) / d;
  }
}

// Wavefront 0 solves the diagonal block: rhs holds b_r - (off-diagonal sums) for lane r,
// col[c] the strictly lower part of column c of the block (lane r holds row r), dval
// the diagonal of M for lane r.  Every lane computes its reciprocal up front; the
// sequential part per column is one corrected multiply, a scalar broadcast and one
// fused update, all in registers (the loop is fully unrolled so the broadcast lane
// is a constant).
template <typename E, typename COL>
__device__ __forceinline__ E solve_diag_block(E rhs, const COL &col, E dval, int unit, int nrows) {
  const int lane = threadIdx.x & 63;
  const E rec = unit ? ezero<E>() : erecip(dval);
  E xr = ezero<E>();
#pragma unroll
  for (int c = 0; c < NB; ++c) {
    if (c < nrows) {
      const E xc_own = unit ? rhs : qdiv(rhs, dval, rec);
      const E xc = rlane(xc_own, c);
      if (lane == c) xr = xc;
      if (lane > c) rhs = emad(rhs, eneg(col(c)), xc);
    }
  }
  return xr;
}

// Published solution blocks.  Every solved x value is also written, with relaxed
// agent-scope atomic stores, to a staging vector that was filled with a signalling-NaN
// sentinel before the launch.  Arithmetic never produces a signalling NaN (IEEE
// operations quiet NaNs), and every published value is the result of arithmetic, so
// a staged word is either the sentinel or its final value: the data is its own ready
// flag.  A consumer polls the 64 values of a block with relaxed atomic loads until no
// lane sees the sentinel; no separate flag, fence or second round trip is needed.
template <typename E> struct Stage;
template <> struct Stage<float> { using W = uint32_t; static constexpr int NW = 1; static constexpr W kSentinel = 0x7FA5A5A5u; };
template <> struct Stage<double> { using W = uint64_t; static constexpr int NW = 1; static constexpr W kSentinel = 0x7FF5A5A5A5A5A5A5ull; };
template <> struct Stage<float2> { using W = uint32_t; static constexpr int NW = 2; static constexpr W kSentinel = 0x7FA5A5A5u; };
template <> struct Stage<double2> { using W = uint64_t; static constexpr int NW = 2; static constexpr W kSentinel = 0x7FF5A5A5A5A5A5A5ull; };

template <typename E>
__device__ __forceinline__ bool stage_load(const typename Stage<E>::W *st, int64_t i, E &out) {
  using S = Stage<E>;
  typename S::W w[S::NW];
  bool ok = true;
#pragma unroll
  for (int k = 0; k < S::NW; ++k) {
    w[k] = __hip_atomic_load(&st[i * S::NW + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    ok = ok && w[k] != S::kSentinel;
  }
  __builtin_memcpy(&out, &w[0], sizeof(E));
  return ok;
}
// The staged words of element i with no check (the check is stage_final, made only
// where the value is needed, so the load is not waited for where it is issued).
template <typename E>
__device__ __forceinline__ E stage_load_raw(const typename Stage<E>::W *st, int64_t i) {
  using S = Stage<E>;
  typename S::W w[S::NW];
#pragma unroll
  for (int k = 0; k < S::NW; ++k) w[k] = __hip_atomic_load(&st[i * S::NW + k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  E out;
  __builtin_memcpy(&out, &w[0], sizeof(E));
  return out;
}
template <typename E>
__device__ __forceinline__ bool stage_final(const E &v) {
  using S = Stage<E>;
  typename S::W w[S::NW];
  __builtin_memcpy(&w[0], &v, sizeof(E));
  bool ok = true;
#pragma unroll
  for (int k = 0; k < S::NW; ++k) ok = ok && w[k] != S::kSentinel;
  return ok;
}

template <typename E>
__device__ __forceinline__ void stage_store(typename Stage<E>::W *st, int64_t i, E v) {
  using S = Stage<E>;
  typename S::W w[S::NW];
  __builtin_memcpy(&w[0], &v, sizeof(E));
#pragma unroll
  for (int k = 0; k < S::NW; ++k) __hip_atomic_store(&st[i * S::NW + k], w[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename W>
__global__ void stage_init_kernel(W *st, int64_t nwords, W sentinel, int *ctrl) {
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i0 == 0) ctrl[0] = 0;  // the ticket counter
  for (int64_t i = i0; i < nwords; i += (int64_t)gridDim.x * blockDim.x) st[i] = sentinel;
}

// Sum of a value over the four lanes of each quad (DPP quad_perm, word by word).
template <typename E> __device__ __forceinline__ E quad_sum(E v) {
  constexpr int NW = sizeof(E) / 4;
  uint32_t w[NW], u[NW];
  __builtin_memcpy(w, &v, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) u[i] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[i], 0xB1, 0xF, 0xF, false);
  E o;
  __builtin_memcpy(&o, u, sizeof(E));
  v = eadd(v, o);  // lanes l, l ^ 1
  __builtin_memcpy(w, &v, sizeof(E));
#pragma unroll
  for (int i = 0; i < NW; ++i) u[i] = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)w[i], 0x4E, 0xF, 0xF, false);
  __builtin_memcpy(&o, u, sizeof(E));
  return eadd(v, o);  // and l ^ 2
}

// inv(L) of the 64 x 64 lower-triangular diagonal block, in place.  On entry Td[c][r]
// holds L(r, c) for c < r (zero elsewhere) and dg[r] = L(r, r) (1 for a unit
// diagonal or a row past n); on exit Td[c][r] = inv(L)(r, c) for c <= r.  S is 48 x
// 16 elements of scratch (in T).  In 16 x 16 blocks:
//  1. wavefront w inverts diagonal block w by forward substitution: lane = 4 cj + p
//     holds column cj of it at rows p, p + 4, p + 8, p + 12; row i's dot product is
//     split over the four phases p and summed across the quad, then divided by its
//     diagonal (rows above the column stay zero, and no term of them is formed, as
//     in substitution);
//  2. block rows 1, 2, 3 in turn: X_ij = -X_ii (sum over k = j .. i-1 of L_ik X_kj)
//     for every j < i, the sums first (into S), then the products over X_ii, with a
//     barrier after each (X_ij overwrites L_ij, which the sums of row i read).
// 16 sequential steps and three levels of 16 x 16 products replace a 64-step
// substitution whose every step broadcast a row through the scalar unit.
template <typename E>
__device__ __forceinline__ void invert_lower64(E (*Td)[NB + 1], const E *dg, E (*S)[NB + 1]) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int cj = lane >> 2, p = lane & 3, c = 16 * w + cj;
  E Y[4] = {ezero<E>(), ezero<E>(), ezero<E>(), ezero<E>()};
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    E s = ezero<E>();
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int k = p + 4 * m;
      if (4 * m < i) {  // k < i is possible for some phase
        const E l = Td[16 * w + k][16 * w + i];
        const E term = emul(l, Y[m]);
        s = (k < i && k >= cj) ? eadd(s, term) : s;
      }
    }
    s = quad_sum(s);
    const E d = dg[16 * w + i];
    const E num = esub(i == cj ? eone<E>() : ezero<E>(), s);
    const E y = i >= cj ? qdiv(num, d, erecip(d)) : ezero<E>();
    if (p == (i & 3)) Y[i >> 2] = y;
  }
  // every lane has read its block of L: write the block's inverse over it
#pragma unroll
  for (int m = 0; m < 4; ++m) Td[c][16 * w + p + 4 * m] = Y[m];
  __syncthreads();
  const int r = t & 15, cc = t >> 4;  // one entry (r, cc) of a 16 x 16 block per thread
#pragma unroll
  for (int i = 1; i < 4; ++i) {
    for (int j = 0; j < i; ++j) {  // S_j = sum_k L_ik X_kj
      E acc = ezero<E>();
      for (int k = j; k < i; ++k)
#pragma unroll
        for (int l = 0; l < 16; ++l) acc = emad(acc, Td[16 * k + l][16 * i + r], Td[16 * j + cc][16 * k + l]);
      S[16 * j + cc][r] = acc;
    }
    __syncthreads();
    for (int j = 0; j < i; ++j) {  // X_ij = -X_ii S_j
      E acc = ezero<E>();
#pragma unroll
      for (int l = 0; l < 16; ++l) acc = emad(acc, Td[16 * i + l][16 * i + r], S[16 * j + cc][l]);
      Td[16 * j + cc][16 * i + r] = eneg(acc);
    }
    __syncthreads();
  }
}

// sum over c of row(c) * xs[c] for this lane, in four partial sums added in a fixed
// order.  xs (16-byte aligned, the same for every lane) is read 16 bytes at a time:
// a wavefront keeps at most 15 LDS reads in flight, so the width of the broadcast
// reads decides how many round trips the product waits for.
template <typename E, class ROW>
__device__ __forceinline__ E row_dot(const ROW &row, const E *xs) {
  constexpr int EV = 16 / sizeof(E);
  typedef uint32_t xv4u __attribute__((ext_vector_type(4)));
  E s4[4] = {ezero<E>(), ezero<E>(), ezero<E>(), ezero<E>()};
#pragma unroll
  for (int c0 = 0; c0 < NB; c0 += EV) {
    const xv4u w = *reinterpret_cast<const xv4u *>(xs + c0);
    E xv[EV];
    __builtin_memcpy(xv, &w, 16);
#pragma unroll
    for (int e = 0; e < EV; ++e) s4[(c0 + e) & 3] = emad(s4[(c0 + e) & 3], row(c0 + e), xv[e]);
  }
  return eadd(eadd(s4[0], s4[1]), eadd(s4[2], s4[3]));
}

// Persistent forward solve of the virtual lower-triangular M (see the file comment).
// RC: the layout of M (1: trans == 0, lanes walk rows; 0: lanes walk columns).
// Block b inverts its diagonal block first (it depends on nothing) while its first
// tiles load.  With ND > 0, blocks from kDirect on take their last ND dependencies off
// the workgroup: the block forms P1 = inv(L_bb) M(b, b-1) (and, ND = 2,
// P2 = inv(L_bb) M(b, b-2)) up front (four wavefronts, 16 columns each; wavefront 0
// keeps its row of P1 in registers, P2 stays in LDS), accumulates tiles j < b - ND as
// block j's values are published, forms y = inv(L_bb) r from that residual (again
// over the four wavefronts), and then wavefront 0 alone waits (ND = 2: for block
// b - 2, u = y - P2 x_{b-2}) for block b - 1, x_b = u - P1 x_{b-1}, and publishes x_b:
// after the tile loop no workgroup barrier is crossed.  The direct path (ND = 0, and
// the first kDirect blocks, whose predecessors publish before the products could be
// ready) accumulates every tile j < b over the four wavefronts and multiplies the
// residual by inv(L_bb) in wavefront 0.
template <typename E, int RC, int ND>
__global__ __launch_bounds__(256) void trsv_kernel(TriView<E> v, E *x, int64_t incx, int unit,
                                                   typename Stage<E>::W *stage, int *ctrl, int nb, int *err,
                                                   uint32_t spin_limit) {
  static_assert(ND >= 0 && ND <= 2, "the last one or two dependencies, or none");
  constexpr bool MT = ND > 0;
  __shared__ E T[NB][NB + 1];
  __shared__ E Td[NB][NB + 1];
  __shared__ E Pb[MT ? NB : 1][NB + 1];
  __shared__ __attribute__((aligned(16))) E xs[NB];
  __shared__ E dg[NB];
  __shared__ E part[4][NB];
  __shared__ int bsh;
  const int t = threadIdx.x, lane = t & 63, q = t >> 6;
  if (t == 0) bsh = atomicAdd(&ctrl[0], 1);
  __syncthreads();
  const int b = bsh;
  if (b >= nb) return;
  const int64_t i0 = (int64_t)b * NB;
  // The first kDirect blocks take the direct path: their predecessors publish before
  // the two products with inv(L_bb) could be ready.
  const bool pre = MT && b >= kDirect;     // uniform over the workgroup
  const int nloop = pre ? b - ND : b;      // tiles accumulated by the four wavefronts
  // wavefront 0's right-hand side, read now: after the tile loop it would put a
  // memory round trip on the critical path of every block
  E rhs0 = ezero<E>();
  if (q == 0 && i0 + lane < v.n) rhs0 = x[v.map(i0 + lane) * incx];
  // the first two tiles of the loop (and, MT, the last two tiles) are in flight while
  // the diagonal block is inverted
  E reg0[16], reg1[16], regp1[16], regp2[16];
  if (nloop > 0) load_tile<RC>(v, i0, 0, false, reg0);
  if (nloop > 1) load_tile<RC>(v, i0, NB, false, reg1);
  if (pre) {
    load_tile<RC>(v, i0, (int64_t)(b - 1) * NB, false, regp1);
    if (ND == 2) load_tile<RC>(v, i0, (int64_t)(b - 2) * NB, false, regp2);
  }
  {
    E dreg[16];
    load_tile<RC>(v, i0, i0, true, dreg);
    store_tile<RC>(v, Td, dreg, i0, i0, true);
    if (q == 0) {
      const int64_t i = i0 + lane;
      dg[lane] = (unit || i >= v.n) ? eone<E>() : v.get(i, i);
    }
    __syncthreads();
    invert_lower64(Td, dg, T);
  }
  // Pb := inv(L_bb) Pb, computed by the four wavefronts (16 columns each) and written
  // back over its input
  auto times_inverse = [&]() {
    E m[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) m[k] = ezero<E>();
    for (int l = 0; l < NB; ++l) {
      const E a = Td[l][lane];  // inv(L_bb)(lane, l)
#pragma unroll
      for (int k = 0; k < 16; ++k) m[k] = emad(m[k], a, Pb[16 * q + k][l]);
    }
    __syncthreads();  // every wavefront is done reading the input
#pragma unroll
    for (int k = 0; k < 16; ++k) Pb[16 * q + k][lane] = m[k];
    __syncthreads();
  };
  E prow[MT ? NB : 1];  // wavefront 0, lane r: row r of P1
  if constexpr (MT) {
    if (pre) {
      store_tile<RC>(v, Pb, regp1, i0, (int64_t)(b - 1) * NB, false);  // M(b, b-1)
      __syncthreads();
      times_inverse();
      if (q == 0) {
#pragma unroll
        for (int c = 0; c < NB; ++c) prow[c] = Pb[c][lane];
      }
      if constexpr (ND == 2) {
        __syncthreads();  // wavefront 0 is done reading P1
        store_tile<RC>(v, Pb, regp2, i0, (int64_t)(b - 2) * NB, false);  // M(b, b-2)
        __syncthreads();
        times_inverse();  // Pb = P2
      }
    }
  }
  E acc = ezero<E>();
  // Wavefront 0 keeps the next block's staged values in flight (xnext, loaded right
  // after the current block's arrive, and checked only when they are needed, so the
  // load costs no wait of its own); a block that was already complete then costs no
  // further load latency.
  E xnext = ezero<E>();
  bool xnext_in = false;  // this lane's row of block xnext_j lies inside the matrix
  int xnext_j = -1;
  auto load_block = [&](int j, E &val) -> bool {  // true when every lane holds a final value
    const int64_t c = (int64_t)j * NB + lane;
    bool ok = true;
    val = ezero<E>();
    if (c < v.n) ok = stage_load<E>(stage, c, val);
    return __ballot(!ok) == 0;
  };
  auto prefetch_block = [&](int j) {  // an unconditional load from a clamped index
    const int64_t c = (int64_t)j * NB + lane;
    xnext_in = c < v.n;
    xnext = stage_load_raw<E>(stage, xnext_in ? c : v.n - 1);
    xnext_j = j;
  };
  // Wavefront 0: block j's published values, from xnext or by polling.  Tickets make
  // this wait finite (block j belongs to a workgroup that started earlier and waits
  // only on smaller tickets); the bounded poll is a backstop that records an error in
  // the stream's error word (reported by coral_stream_check and by the host-pointer
  // form) instead of hanging if that invariant were broken.  The default bound is
  // seconds; a legitimate wait is microseconds.
  auto wait_block = [&](int j) -> E {
    E val = xnext_in ? xnext : ezero<E>();
    if (spin_limit == 0) {  // the host's test mode: every wait gives up at once
      if (lane == 0) atomicOr(err, (int)ERRF_WAIT_BOUND);
    } else if (xnext_j != j || __ballot(xnext_in && !stage_final(xnext)) != 0) {
      uint32_t polls = 0;  // failed polls; the bound counts them
      while (!load_block(j, val)) {
        if (++polls >= spin_limit) {
          if (lane == 0) atomicOr(err, (int)ERRF_WAIT_BOUND);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    xnext_j = -1;
    return val;
  };
  // Off-diagonal tiles stream through two register slots (slot j&1 holds tile j,
  // loaded two tiles ahead); the unrolled-by-two loop keeps slot indices static.
  auto step = [&](int j, E(&cur)[16]) {
    __syncthreads();  // every wavefront is done reading T and xs of tile j-1
    store_tile<RC>(v, T, cur, i0, (int64_t)j * NB, false);
    if (q == 0) {
      xs[lane] = wait_block(j);
      if (j + 1 < b) prefetch_block(j + 1);
    }
    __syncthreads();
    if (j + 2 < nloop) load_tile<RC>(v, i0, (int64_t)(j + 2) * NB, false, cur);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int c = q * 16 + k;
      acc = emad(acc, T[c][lane], xs[c]);
    }
  };
  for (int j = 0; j < nloop; j += 2) {
    step(j, reg0);
    if (j + 1 < nloop) step(j + 1, reg1);
  }
  part[q][lane] = acc;
  __syncthreads();
  const int64_t i = i0 + lane;
  // four partial sums in a fixed order keep each chain short and the result
  // deterministic
  if (pre) {
    // the residual of tiles j < b - ND, then y = inv(L_bb) r over the four wavefronts
    if (q == 0) {
      E r = ezero<E>();
      if (i < v.n) r = esub(rhs0, eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane])));
      xs[lane] = r;
    }
    __syncthreads();
    E yq = ezero<E>();
#pragma unroll
    for (int k = 0; k < 16; ++k) yq = emad(yq, Td[16 * q + k][lane], xs[16 * q + k]);
    __syncthreads();  // every wavefront has read xs and the residual's part sums
    part[q][lane] = yq;
    __syncthreads();
    if (q == 0) {
      E xr = eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane]));
      if constexpr (ND == 2) {  // u = y - P2 x_{b-2}, P2 from LDS
        xs[lane] = wait_block(b - 2);
        prefetch_block(b - 1);
        __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wavefront's LDS writes are done
        __builtin_amdgcn_wave_barrier();
        xr = esub(xr, row_dot<E>([&](int c) { return Pb[c][lane]; }, xs));
        __builtin_amdgcn_wave_barrier();  // every lane has read xs
      }
      {  // x_b = u - P1 x_{b-1}, P1 from registers
        xs[lane] = wait_block(b - 1);
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        xr = esub(xr, row_dot<E>([&](int c) { return prow[c]; }, xs));
      }
      if (i < v.n) {
        stage_store<E>(stage, i, xr);  // publish first: it is the critical path
        x[v.map(i) * incx] = xr;
      }
    }
  } else {
    if (q == 0) {
      E rhs = ezero<E>();
      if (i < v.n) rhs = esub(rhs0, eadd(eadd(part[0][lane], part[1][lane]), eadd(part[2][lane], part[3][lane])));
      // x = inv(L_bb) r: lane i forms row i from the broadcast residual
      xs[lane] = rhs;
      __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0): this wavefront's LDS writes are done
      __builtin_amdgcn_wave_barrier();
      const E xr = row_dot<E>([&](int c) { return Td[c][lane]; }, xs);
      if (i < v.n) {
        stage_store<E>(stage, i, xr);  // publish first: it is the critical path
        x[v.map(i) * incx] = xr;
      }
    }
  }
}

__global__ void zero_ints_kernel(int *p, int count) {
  for (int i = threadIdx.x + blockIdx.x * blockDim.x; i < count; i += blockDim.x * gridDim.x) p[i] = 0;
}

// CORAL argument order: uplo=1 trans=2 diag=3 n=4 a=5 lda=6 x=7 incx=8
template <typename E>
static int check_tr(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const E *a, int64_t lda, E *x,
                    int64_t incx) {
  if (!valid_uplo(uplo)) return -1;
  if (!valid_op(trans)) return -2;
  if (!valid_diag(diag)) return -3;
  if (n < 0) return -4;
  if (lda < imax64(1, n)) return -6;
  if (incx <= 0) return -8;
  if (n > 0 && !a) return -5;
  if (n > 0 && !x) return -7;
  return CORAL_OK;
}

template <typename E>
static TriView<E> make_view(int32_t uplo, int32_t trans, const E *a, int64_t lda, int64_t n) {
  TriView<E> v;
  v.A = a;
  v.lda = lda;
  v.n = n;
  // real routines map ConjugateTranspose to Transpose (coral-aarch64/src/level2/dtrlsv.rs:433)
  v.trans = (trans == CORAL_CONJUGATE_TRANSPOSE && !elem_traits<E>::cplx) ? 1 : trans;
  const bool lower = uplo == CORAL_LOWER_TRIANGULAR;
  v.fwd = (lower == (trans == CORAL_NO_TRANSPOSE)) ? 1 : 0;
  return v;
}

// How many of a block's last dependencies trsv_kernel precomputes (ND), per element
// type.  Complex double has no LDS for the products (0).  Measured on one box, lower
// non-unit n = 8192, N / T, us, for ND = 2 / 1 / 0: STRSV 183 / 138, 178 / 137, 193 / 176;
// DTRSV 238 / 211, 221 / 197, 223 / 209; CTRSV 326 / 330, 259 / 238, 260 / 293.
// CORAL_TRSV_DEPS=<0|1|2> overrides it (complex double stays at 0).
template <typename E> static int trsv_deps() {
  int nd = sizeof(E) > 8 ? 0 : 1;
  if (const char *e = getenv("CORAL_TRSV_DEPS")) nd = atoi(e);
  return nd < 0 ? 0 : (nd > 2 ? 2 : nd);
}

template <typename E>
static int trsv_impl(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const E *a, int64_t lda, E *x,
                     int64_t incx, coral_stream_t stream) {
  CORAL_TRY(check_tr(uplo, trans, diag, n, a, lda, x, incx));
  if (n == 0) return CORAL_OK;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  const int nb = (int)cdiv(n, NB);
  using W = typename Stage<E>::W;
  void *ctrl = nullptr, *stage = nullptr;
  int *err = nullptr;
  // ctrl: [0] ticket; stage: published x, sentinel-filled; err: the stream's error word
  CORAL_TRY(workspace(s, WS_SMALL, 2 * sizeof(int), &ctrl));
  CORAL_TRY(error_word(s, &err));
  const int64_t nwords = n * Stage<E>::NW;
  CORAL_TRY(workspace(s, WS_AUX2, sizeof(W) * (size_t)nwords, &stage));
  hipLaunchKernelGGL(stage_init_kernel<W>, dim3((unsigned)std::min<int64_t>(cdiv(nwords, 256), 1024)), dim3(256), 0, s,
                     (W *)stage, nwords, Stage<E>::kSentinel, (int *)ctrl);
  CORAL_LAUNCH_CHECK();
  TriView<E> v = make_view(uplo, trans, a, lda, n);
  const int nd = trsv_deps<E>();
  const int udiag = (int)(diag == CORAL_UNIT_DIAGONAL);
  const uint32_t lim = wait_limit();
#define CORAL_TRSV_LAUNCH(RC, ND)                                                                             \
  hipLaunchKernelGGL((trsv_kernel<E, RC, ND>), dim3((unsigned)nb), dim3(256), 0, s, v, x, incx, udiag, (W *)stage, \
                     (int *)ctrl, nb, err, lim)
  if constexpr (sizeof(E) > 8) {
    if (v.trans == 0) CORAL_TRSV_LAUNCH(1, 0);
    else CORAL_TRSV_LAUNCH(0, 0);
  } else {
    if (v.trans == 0) {
      if (nd == 2) CORAL_TRSV_LAUNCH(1, 2);
      else if (nd == 1) CORAL_TRSV_LAUNCH(1, 1);
      else CORAL_TRSV_LAUNCH(1, 0);
    } else {
      if (nd == 2) CORAL_TRSV_LAUNCH(0, 2);
      else if (nd == 1) CORAL_TRSV_LAUNCH(0, 1);
      else CORAL_TRSV_LAUNCH(0, 0);
    }
  }
#undef CORAL_TRSV_LAUNCH
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// ---- batched DTRSV: persistent workgroups pull whole problems from a queue ----
// The working solution of a problem lives in LDS when it fits (n <= kBatchMaxN);
// larger problems work on x in global memory (the workgroup's own writes, made
// visible to its other waves by the barrier after each diagonal block).  A problem
// with n < 0 or lda < max(1, n) is skipped and flagged in the stream's error word.
constexpr int kBatchMaxN = 4096;  // 32 KiB of doubles in LDS

// One problem's solve by one workgroup.  IN_LDS: the working x is xsol (LDS), else
// x itself in global memory; RC: the layout of M (1: trans == 0), so that the tile
// loads are the unconditional clamped ones.  Templates, so that no branch on either
// sits in the inner loops: a branch join after a global load makes the compiler wait
// for every load in flight.
template <bool IN_LDS, int RC>
__device__ __forceinline__ void batched_solve(const TriView<double> &v, double *x, int64_t n, int unit,
                                              double (&T)[NB][NB + 1], double *xsol, double (&part)[4][NB]) {
  const int t = threadIdx.x, lane = t & 63, q = t >> 6;
  auto xget = [&](int64_t i) -> double { return IN_LDS ? xsol[i] : x[v.map(i)]; };
  if (IN_LDS) {
    for (int64_t i = t; i < n; i += 256) xsol[i] = x[v.map(i)];
    __syncthreads();
  }
  const int nbp = (int)cdiv(n, NB);
  double reg[16];
  for (int b = 0; b < nbp; ++b) {
    const int64_t i0 = (int64_t)b * NB;
    double acc = 0.0;
    for (int j = 0; j < b; ++j) {
      load_tile<RC>(v, i0, (int64_t)j * NB, false, reg);
      store_tile<RC>(v, T, reg, i0, (int64_t)j * NB, false);
      __syncthreads();
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        const int c = q * 16 + k;
        acc = fma(T[c][lane], xget((int64_t)j * NB + c), acc);  // j * NB + c < i0 <= n: solved
      }
      __syncthreads();
    }
    part[q][lane] = acc;
    load_tile<RC>(v, i0, i0, true, reg);
    store_tile<RC>(v, T, reg, i0, i0, true);
    __syncthreads();
    if (q == 0) {
      const int64_t i = i0 + lane;
      const int nrows = (int)std::min<int64_t>(NB, n - i0);
      double rhs = 0.0, dval = 0.0;
      if (i < n) {
        rhs = xget(i) - ((part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]));
        if (!unit) dval = v.get(i, i);
      }
      double xr = solve_diag_block(rhs, [&](int c) { return T[c][lane]; }, dval, unit, nrows);
      if (i < n) {
        if (IN_LDS) xsol[i] = xr;
        else x[v.map(i)] = xr;
      }
    }
    __syncthreads();
  }
  if (IN_LDS)
    for (int64_t i = t; i < n; i += 256) x[v.map(i)] = xsol[i];
  __syncthreads();
}

// Two workgroups per CU (the LDS allows two): the bound keeps the registers within
// two waves per SIMD.
__global__ __launch_bounds__(256, 2) void trsv_batched_kernel(int uplo, int trans, int unit, int64_t batch,
                                                           const int32_t *ns, const double *const *As,
                                                           const int32_t *ldas, double *const *xs_g, int *ctrl,
                                                           int *err) {
  __shared__ double T[NB][NB + 1];
  __shared__ double xsol[kBatchMaxN];
  __shared__ double part[4][NB];
  __shared__ int psh;
  const int t = threadIdx.x;
  for (;;) {
    if (t == 0) psh = atomicAdd(&ctrl[0], 1);
    __syncthreads();
    const int64_t p = psh;
    __syncthreads();
    if (p >= batch) return;  // every wavefront of the workgroup leaves together
    const int64_t n = ns[p];
    const int64_t lda = ldas[p];
    if (n == 0) continue;
    if (n < 0 || lda < n) {
      if (t == 0) atomicOr(err, (int)ERRF_PROBLEM_ARG);
      continue;
    }
    TriView<double> v;
    v.A = As[p];
    v.lda = lda;
    v.n = n;
    v.trans = trans == CORAL_CONJUGATE_TRANSPOSE ? 1 : trans;
    const bool lower = uplo == CORAL_LOWER_TRIANGULAR;
    v.fwd = (lower == (trans == CORAL_NO_TRANSPOSE)) ? 1 : 0;
    // uniform branches: n and the layout are the same for the whole workgroup
    if (n <= kBatchMaxN) {
      if (v.trans == 0) batched_solve<true, 1>(v, xs_g[p], n, unit, T, xsol, part);
      else batched_solve<true, 0>(v, xs_g[p], n, unit, T, xsol, part);
    } else {
      if (v.trans == 0) batched_solve<false, 1>(v, xs_g[p], n, unit, T, xsol, part);
      else batched_solve<false, 0>(v, xs_g[p], n, unit, T, xsol, part);
    }
  }
}

}  // namespace tri
}  // namespace coral

using namespace coral;
using namespace coral::tri;

extern "C" {

#define TRSV_ENTRY(NAME, E, T)                                                                               \
  int coral_##NAME(int32_t uplo, int32_t trans, int32_t diag, int64_t n, const T *a, int64_t lda, T *x,     \
                   int64_t incx, coral_stream_t s) {                                                         \
    return trsv_impl<E>(uplo, trans, diag, n, (const E *)a, lda, (E *)x, incx, s);                          \
  }
TRSV_ENTRY(strsv, float, float)
TRSV_ENTRY(dtrsv, double, double)
TRSV_ENTRY(ctrsv, float2, float)
TRSV_ENTRY(ztrsv, double2, double)

// uplo=1 trans=2 diag=3 batch=4 n=5 a=6 lda=7 x=8
int coral_dtrsv_batched(int32_t uplo, int32_t trans, int32_t diag, int64_t batch, const int32_t *n,
                        const double *const *a, const int32_t *lda, double *const *x, coral_stream_t stream) {
  if (!valid_uplo(uplo)) return -1;
  if (!valid_op(trans)) return -2;
  if (!valid_diag(diag)) return -3;
  if (batch < 0 || batch > ((int64_t)1 << 30)) return -4;
  if (batch == 0) return CORAL_OK;
  if (!n) return -5;
  if (!a) return -6;
  if (!lda) return -7;
  if (!x) return -8;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  void *ctrl = nullptr;
  int *err = nullptr;
  CORAL_TRY(workspace(s, WS_SMALL, 64, &ctrl));
  CORAL_TRY(error_word(s, &err));
  hipLaunchKernelGGL(zero_ints_kernel, dim3(1), dim3(64), 0, s, (int *)ctrl, 16);
  CORAL_LAUNCH_CHECK();
  const int64_t grid = std::min<int64_t>(batch, 2 * (int64_t)num_cus());
  hipLaunchKernelGGL(trsv_batched_kernel, dim3((unsigned)grid), dim3(256), 0, s, uplo, trans,
                     (int)(diag == CORAL_UNIT_DIAGONAL), batch, n, a, lda, x, (int *)ctrl, err);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

}  // extern "C"
