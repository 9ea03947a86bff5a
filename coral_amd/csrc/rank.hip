// Level-2 rank updates for gfx950: ?GER / ?GERU / ?GERC, ?SYR / ?SYR2, ?HER / ?HER2.
//
// CORAL runs these column by column: a coefficient per column (alpha*y_j, alpha*x_j
// or their conjugate forms), skipped when it is zero, then one or two axpys of x
// (and y) into the stored rows of that column (coral-aarch64/src/level2/
// sger.rs:68-132, cgeru.rs:63-144, cgerc.rs:62-143, ssyr.rs:63-181,
// ssyr2.rs:59-197, cher.rs:62-212, cher2.rs:65-311 and the d/z twins).
//
// Here one launch streams the stored part of A once, read-modify-write: a
// workgroup owns 256*EV rows x 4, 8 or 16 columns (by kind; EV elements per 16-byte vector),
// keeps its rows of x (and y) in registers and has four columns of 16-byte
// nontemporal loads in flight per thread.  Every wave instruction covers one
// contiguous 1 KiB run of a column.
//
// The results are bit-identical to CORAL's: each element receives the same
// operations in the same order.  CORAL's unit-stride axpy runs a NEON body with
// fused multiply-adds over the first floor(len/Q)*Q elements of a segment and a
// scalar tail `y += c*x` with two roundings (Q = 8 saxpy, 4 daxpy and caxpy,
// 2 zaxpy: saxpy.rs:63-112, daxpy.rs:63-112, caxpy.rs:70-146, zaxpy.rs:69-145);
// its strided axpy has two roundings throughout; the strided real rank updates
// use `x.mul_add(c, a)` (fused) inline.  The element updates are axpy_upd
// (coral_common.hpp); this file is compiled with -ffp-contract=off so that the
// per-column coefficients also keep CORAL's separate roundings.
#include "coral_common.hpp"

#include <algorithm>
#include <type_traits>

namespace coral {
namespace rk {

typedef uint32_t rv4u __attribute__((ext_vector_type(4)));

enum Kind { K_GER = 0, K_GERU = 1, K_GERC = 2, K_SYR = 3, K_SYR2 = 4, K_HER = 5, K_HER2 = 6 };
// How an element update rounds (see the file comment).
enum Mode { M_PLAIN = 0, M_UNIT = 1, M_FUSED = 2 };

constexpr int kThreads = 256;
constexpr int kU = 4;      // columns in flight per thread

template <typename E>
struct RankArgs {
  int64_t m, n;       // rows and columns of A (m == n for SYR/SYR2/HER/HER2)
  E alpha;            // HER: the real alpha in .x
  const E *x;
  int64_t incx;
  const E *y;
  int64_t incy;
  E *a;
  int64_t lda;
  int64_t cpc;        // columns per workgroup
  int lower;          // stored triangle (triangular kinds)
  int mode1, mode2;   // rounding mode of the x update / the y update
  int unit;           // incx == 1 && incy == 1
  int vec_a;          // A and lda allow aligned 16-byte column vectors
  int vec_x, vec_y;   // x / y are unit stride and 16-byte aligned
};

// number of leading elements of a segment of length len that get the fused update
template <typename E>
__device__ __forceinline__ int64_t fused_prefix(int mode, int64_t len) {
  constexpr int64_t Q = axpy_body<E>::Q;
  return mode == M_FUSED ? len : (mode == M_UNIT ? len / Q * Q : 0);
}

template <typename E> __device__ __forceinline__ bool nz(E v) { return !eiszero(v); }

// Per-column coefficients (c1 multiplies x, c2 multiplies y).
template <typename E, int K>
__device__ __forceinline__ void column_coefs(const RankArgs<E> &p, int64_t j, E &c1, E &c2) {
  c1 = ezero<E>();
  c2 = ezero<E>();
  if constexpr (K == K_GER) {  // sger.rs:95 / :117: coeff = alpha * y_j
    c1 = p.alpha * p.y[j * p.incy];
  } else if constexpr (K == K_SYR) {  // ssyr.rs:92 / :143: aj = alpha * x_j
    c1 = p.alpha * p.x[j * p.incx];
  } else if constexpr (K == K_SYR2) {  // ssyr2.rs:88-89 / :155-156: aj_y = alpha*y_j, aj_x = alpha*x_j
    c1 = p.alpha * p.y[j * p.incy];
    c2 = p.alpha * p.x[j * p.incx];
  } else if constexpr (K == K_GERU || K == K_GERC) {
    const E yv = p.y[j * p.incy];
    const auto ar = p.alpha.x, ai = p.alpha.y;
    if constexpr (K == K_GERU) {  // cgeru.rs:92-95 / :120-121
      c1.x = ar * yv.x - ai * yv.y;
      c1.y = ar * yv.y + ai * yv.x;
    } else {  // cgerc.rs:91-94 / :119-120
      c1.x = ar * yv.x + ai * yv.y;
      c1.y = -ar * yv.y + ai * yv.x;
    }
  } else if constexpr (K == K_HER) {  // cher.rs:93: aj = [alpha*xr, -alpha*xi]
    const E xv = p.x[j * p.incx];
    c1.x = p.alpha.x * xv.x;
    c1.y = -p.alpha.x * xv.y;
  } else {  // K_HER2, cher2.rs:106-109
    const E xv = p.x[j * p.incx], yv = p.y[j * p.incy];
    const auto ar = p.alpha.x, ai = p.alpha.y;
    c1.x = ar * yv.x + ai * yv.y;
    c1.y = -ar * yv.y + ai * yv.x;
    c2.x = ar * xv.x - ai * xv.y;
    c2.y = -ar * xv.y - ai * xv.x;
  }
}

// Which updates a column gets (CORAL's skip rules): ap1 the x update, ap2 the y
// update; zdiag: Im(a_jj) is set to zero after them (HER/HER2).
template <typename E, int K>
__device__ __forceinline__ void column_flags(const RankArgs<E> &p, E c1, E c2, bool &ap1, bool &ap2, bool &zdiag) {
  ap1 = nz(c1);
  ap2 = false;
  zdiag = false;
  if constexpr (K == K_SYR2) {
    if (p.unit) {
      ap2 = nz(c2);
    } else {  // one mul_add loop with both coefficients unless both are zero (ssyr2.rs:157-168)
      ap1 = ap2 = nz(c1) || nz(c2);
    }
  } else if constexpr (K == K_HER) {
    zdiag = ap1;
  } else if constexpr (K == K_HER2) {
    ap2 = nz(c2);
    // unit strides zero the diagonal's imaginary part for every column
    // (cher2.rs:136), strided calls only for updated columns (cher2.rs:215-244)
    zdiag = p.unit || ap1 || ap2;
  }
}

// Rows [r, r+EV) of a source vector into registers (zero beyond m).
template <typename E>
__device__ __forceinline__ void load_src(E (&v)[16 / sizeof(E)], const E *s, int64_t inc, bool vec, int64_t r,
                                         int64_t m) {
  constexpr int EV = 16 / sizeof(E);
  if (vec && r + EV <= m) {
    const rv4u pk = *reinterpret_cast<const rv4u *>(s + r);
    __builtin_memcpy(&v[0], &pk, 16);
  } else {
#pragma unroll
    for (int e = 0; e < EV; ++e) v[e] = (r + e < m) ? s[(r + e) * inc] : ezero<E>();
  }
}

// v as held by lane l of the wavefront (l wave-uniform), word by word.
template <typename E>
__device__ __forceinline__ E wave_bcast(E v, int l) {
  constexpr int W = sizeof(E) / 4;
  uint32_t u[W];
  __builtin_memcpy(u, &v, sizeof(E));
#pragma unroll
  for (int k = 0; k < W; ++k) u[k] = (uint32_t)__builtin_amdgcn_readlane((int)u[k], l);
  __builtin_memcpy(&v, u, sizeof(E));
  return v;
}

// Largest g with g(g+1)/2 <= q.
__device__ __forceinline__ int64_t tri_root(int64_t q) {
  int64_t g = (int64_t)((sqrt(8.0 * (double)q + 1.0) - 1.0) * 0.5);
  while (g * (g + 1) / 2 > q) --g;
  while ((g + 1) * (g + 2) / 2 <= q) ++g;
  return g;
}

// Rows [l, h) of column j that the update touches (the stored part), and the
// start of the axpy segment CORAL runs over them.
template <int K>
__device__ __forceinline__ void column_rows(int64_t m, int64_t n, int lower, int64_t j, int64_t &l, int64_t &h) {
  constexpr bool TRI = K == K_SYR || K == K_SYR2 || K == K_HER || K == K_HER2;
  l = 0;
  h = m;
  if (TRI) {
    l = lower ? j : 0;
    h = lower ? n : j + 1;
  }
}

// One workgroup: rows [r0, r0 + 256*EV) x columns [c0, c0 + cpc).  GER uses a 2-D
// grid, row block fastest, so that consecutive workgroups walk down the same
// columns.  The triangular kinds enumerate only the tiles that meet the stored
// triangle, in a 1-D grid: column blocks are grouped KT = 256*EV/cpc at a time
// (one row block's width); the groups are ordered so that the k-th meets k+1 row
// blocks (upper: group k, row blocks 0..k; lower: group nrb-1-k, its last k+1 row
// blocks), so tile id falls in group tri_root(id/KT), and inside a group the row
// block runs fastest.  Measured on 16384^2 DSYR: a 2-D grid with empty
// workgroups ran at 2.97 TB/s, this enumeration at 5.0-5.7 TB/s.
template <typename E, int K>
__global__ __launch_bounds__(kThreads) void rank_kernel(RankArgs<E> p) {
  constexpr int EV = 16 / sizeof(E);
  constexpr int64_t RB = kThreads * EV;
  constexpr bool TRI = K == K_SYR || K == K_SYR2 || K == K_HER || K == K_HER2;
  constexpr bool TWO = K == K_SYR2 || K == K_HER2;
  int64_t rb = blockIdx.x, cb = blockIdx.y;
  if (TRI) {
    const int64_t KT = RB / p.cpc;
    const int64_t id = blockIdx.x;
    const int64_t g = tri_root(id / KT);
    const int64_t w0 = id - KT * (g * (g + 1) / 2);
    // tiles of one group of KT column blocks: g+1 row blocks, row block fastest
    const int64_t grp = p.lower ? cdiv(p.m, RB) - 1 - g : g;
    rb = (p.lower ? grp : 0) + w0 % (g + 1);
    cb = KT * grp + w0 / (g + 1);
    if (cb * p.cpc >= p.n || rb * RB >= p.m) return;
  }
  const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t r0 = rb * RB;
  const int64_t c0 = cb * p.cpc;
  const int64_t c1 = std::min<int64_t>(c0 + p.cpc, p.n);
  const int64_t rw = r0 + (int64_t)w * 64 * EV;  // first row of this wave
  if (rw >= p.m) return;                          // uniform per wave; no barrier below
  const int64_t rr = rw + (int64_t)lane * EV;     // first row of this thread

  E xs[EV], ys[EV];
  load_src<E>(xs, p.x, p.incx, p.vec_x, rr, p.m);
  if (TWO) load_src<E>(ys, p.y, p.incy, p.vec_y, rr, p.m);

  // Columns in groups of 64: lane t computes the coefficients of column g0 + t once,
  // and the column loop takes them from that lane.  (Loaded per column inside the
  // loop, each coefficient is a load that the next update waits for, together with
  // every load of A issued before it.)  Real double is the exception: its columns
  // stream fastest with the coefficient loaded beside each column and one column in
  // flight per wave.  Measured at 16384^2 / 8192^2, two runs each: DGER 6.02-6.04
  // TB/s that way, 5.30-5.41 with broadcast coefficients and the unguarded path
  // below, 5.56-5.83 with broadcast coefficients alone; the other element types gain
  // from both (SGER 5.60 -> 6.26-6.28, CGERU 5.53-5.57 -> 6.15-6.17, ZGERC
  // 5.49-5.58 -> 5.78-5.81, CHER 4.90-4.93 -> 5.01-5.02, ZHER2 4.55-4.60 -> 5.21-5.24).
  constexpr bool kWide = !std::is_same<E, double>::value;
  for (int64_t g0 = c0; g0 < c1; g0 += 64) {
    const int64_t g1 = std::min<int64_t>(g0 + 64, c1);
    E lc1 = ezero<E>(), lc2 = ezero<E>();
    if (kWide && g0 + lane < g1) column_coefs<E, K>(p, g0 + lane, lc1, lc2);
    for (int64_t j0 = g0; j0 < g1; j0 += kU) {
      E av[kU][EV];
      E cf1[kU], cf2[kU];
      // Fast path (wave-uniform): all kU columns exist and this wave's rows lie inside
      // every one's stored rows.  Loads and stores are then unconditional: a guarded
      // load is a branch, and at each branch join the compiler waits for every load
      // issued so far, which would leave one column in flight instead of kU.  A
      // column CORAL skips (zero coefficient) is read and written back unchanged
      // here, which leaves the same bits.
      bool fast = kWide && p.vec_a && j0 + kU <= g1;
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        int64_t l, h;
        column_rows<K>(p.m, p.n, p.lower, j0 + u, l, h);
        fast = fast && rw >= l && rw + 64 * EV <= h;
      }
      if (fast) {
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const rv4u pk = __builtin_nontemporal_load(reinterpret_cast<const rv4u *>(p.a + (j0 + u) * p.lda + rr));
          __builtin_memcpy(&av[u][0], &pk, 16);
        }
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          const int64_t j = j0 + u;
          cf1[u] = wave_bcast(lc1, (int)(j - g0));
          cf2[u] = TWO ? wave_bcast(lc2, (int)(j - g0)) : ezero<E>();
          bool a1, a2, zd;
          column_flags<E, K>(p, cf1[u], cf2[u], a1, a2, zd);
          int64_t l, h;
          column_rows<K>(p.m, p.n, p.lower, j, l, h);
          const int64_t nf1 = fused_prefix<E>(p.mode1, h - l), nf2 = fused_prefix<E>(p.mode2, h - l);
#pragma unroll
          for (int e = 0; e < EV; ++e) {
            const int64_t i = rr + e, t = i - l;
            E v = av[u][e];
            if (a1) v = axpy_upd(v, cf1[u], xs[e], t < nf1);
            if (TWO && a2) v = axpy_upd(v, cf2[u], ys[e], t < nf2);
            if constexpr (elem_traits<E>::cplx) {
              if (zd && i == j) v.y = 0;
            }
            av[u][e] = v;
          }
        }
        // stores after all updates: a store issued between them would be waited on
        // at the next branch join
#pragma unroll
        for (int u = 0; u < kU; ++u) {
          rv4u pk;
          __builtin_memcpy(&pk, &av[u][0], 16);
          __builtin_nontemporal_store(pk, reinterpret_cast<rv4u *>(p.a + (j0 + u) * p.lda + rr));
        }
        continue;
      }
      // loads of kU columns first (coefficients and flags are wave-uniform)
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t j = j0 + u;
        if (j >= g1) continue;
        if constexpr (kWide) {
          cf1[u] = wave_bcast(lc1, (int)(j - g0));
          cf2[u] = TWO ? wave_bcast(lc2, (int)(j - g0)) : ezero<E>();
        } else {
          column_coefs<E, K>(p, j, cf1[u], cf2[u]);
        }
        bool a1, a2, zd;
        column_flags<E, K>(p, cf1[u], cf2[u], a1, a2, zd);
        int64_t l, h;
        column_rows<K>(p.m, p.n, p.lower, j, l, h);
        if (!(a1 || a2)) {  // at most the diagonal changes (HER2 with zero coefficients)
          if (!zd) continue;
          l = j;
          h = j + 1;
        }
        if (rw >= h || rw + 64 * EV <= l) continue;  // no row of this wave
        const E *col = p.a + j * p.lda;
        if (p.vec_a && rr >= l && rr + EV <= h) {
          const rv4u pk = __builtin_nontemporal_load(reinterpret_cast<const rv4u *>(col + rr));
          __builtin_memcpy(&av[u][0], &pk, 16);
        } else {
#pragma unroll
          for (int e = 0; e < EV; ++e) av[u][e] = (rr + e >= l && rr + e < h) ? col[rr + e] : ezero<E>();
        }
      }
      // then the updates and stores; the per-column state is recomputed from j
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t j = j0 + u;
        if (j >= g1) continue;
        bool a1, a2, zd;
        column_flags<E, K>(p, cf1[u], cf2[u], a1, a2, zd);
        int64_t l, h;
        column_rows<K>(p.m, p.n, p.lower, j, l, h);
        const int64_t seg0 = l;
        const int64_t nf1 = fused_prefix<E>(p.mode1, h - l), nf2 = fused_prefix<E>(p.mode2, h - l);
        if (!(a1 || a2)) {
          if (!zd) continue;
          l = j;
          h = j + 1;
        }
        if (rw >= h || rw + 64 * EV <= l) continue;
#pragma unroll
        for (int e = 0; e < EV; ++e) {
          const int64_t i = rr + e, t = i - seg0;
          E v = av[u][e];
          if (a1) v = axpy_upd(v, cf1[u], xs[e], t < nf1);
          if (TWO && a2) v = axpy_upd(v, cf2[u], ys[e], t < nf2);
          if constexpr (elem_traits<E>::cplx) {
            if (zd && i == j) v.y = 0;
          }
          av[u][e] = v;
        }
        E *col = p.a + j * p.lda;
        if (p.vec_a && rr >= l && rr + EV <= h) {
          rv4u pk;
          __builtin_memcpy(&pk, &av[u][0], 16);
          __builtin_nontemporal_store(pk, reinterpret_cast<rv4u *>(col + rr));
        } else {
#pragma unroll
          for (int e = 0; e < EV; ++e)
            if (rr + e >= l && rr + e < h) col[rr + e] = av[u][e];
        }
      }
    }
  }
}

// CORAL argument positions of the checked arguments, per entry point.
struct ArgPos { int uplo, m, n, alpha, x, incx, y, incy, a, lda; };

template <typename E, int K>
int rank_impl(const ArgPos &ap, int32_t uplo, int64_t m, int64_t n, E alpha, const E *x, int64_t incx, const E *y,
              int64_t incy, E *a, int64_t lda, coral_stream_t stream) {
  constexpr bool TRI = K == K_SYR || K == K_SYR2 || K == K_HER || K == K_HER2;
  constexpr bool HASY = !(K == K_SYR || K == K_HER);
  if (TRI && !valid_uplo(uplo)) return -ap.uplo;
  if (TRI) m = n;
  if (m < 0) return -ap.m;
  if (n < 0) return -ap.n;
  if (incx <= 0) return -ap.incx;
  if (HASY && incy <= 0) return -ap.incy;
  if (lda < imax64(1, m)) return -ap.lda;
  // quick returns: sger.rs:80, ssyr.rs:73, ssyr2.rs:70, cher.rs:72, cher2.rs:77
  if (m == 0 || n == 0 || eiszero(alpha)) return CORAL_OK;
  if (!x) return -ap.x;
  if (HASY && !y) return -ap.y;
  if (!a) return -ap.a;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);

  RankArgs<E> p;
  p.m = m;
  p.n = n;
  p.alpha = alpha;
  p.x = x;
  p.incx = incx;
  p.y = HASY ? y : x;
  p.incy = HASY ? incy : incx;
  p.a = a;
  p.lda = lda;
  p.lower = uplo == CORAL_LOWER_TRIANGULAR;
  const bool unit = incx == 1 && (!HASY || incy == 1);
  p.unit = unit;
  // rounding modes, per CORAL's dispatch on the strides
  switch (K) {
    case K_GER: p.mode1 = p.mode2 = unit ? M_UNIT : M_FUSED; break;   // sger.rs:92 / :125
    case K_GERU:
    case K_GERC: p.mode1 = p.mode2 = unit ? M_UNIT : M_PLAIN; break;  // cgerc.rs:86 / :131-135
    case K_SYR: p.mode1 = p.mode2 = incx == 1 ? M_UNIT : M_FUSED; break;  // ssyr.rs:86 / :152
    case K_SYR2: p.mode1 = p.mode2 = unit ? M_UNIT : M_FUSED; break;      // ssyr2.rs:84 / :166-167
    case K_HER: p.mode1 = p.mode2 = incx == 1 ? M_UNIT : M_PLAIN; break;  // cher.rs:85: caxpy(x, incx, col, 1)
    default:  // K_HER2: caxpy(x, incx, col, 1) then caxpy(y, incy, col, 1)
      p.mode1 = incx == 1 ? M_UNIT : M_PLAIN;
      p.mode2 = incy == 1 ? M_UNIT : M_PLAIN;
  }
  p.vec_a = ((uintptr_t)a % 16 == 0) && ((lda * (int64_t)sizeof(E)) % 16 == 0);
  p.vec_x = incx == 1 && ((uintptr_t)x % 16 == 0);
  p.vec_y = p.incy == 1 && ((uintptr_t)p.y % 16 == 0);

  constexpr int EV = 16 / sizeof(E);
  constexpr int64_t RB = (int64_t)kThreads * EV;
  const int64_t row_blocks = cdiv(m, RB);
  // Columns per workgroup, per kind.  Measured (TB/s, 16384^2 real / 8192^2 complex,
  // columns 16 -> 8 -> 4): SGER 6.28 -> 6.33 -> 6.61, DGER 6.01 -> 6.21 -> 6.20,
  // CGERU 6.16 -> 6.24 -> 6.60, ZGERC 5.76 -> 6.16 -> 6.34, CHER 4.95 -> 5.92 -> 6.19,
  // DSYR 5.61 -> 5.73 -> 5.32, ZHER2 5.18 -> 5.41-5.50 -> 5.03, DSYR2 5.26 -> 5.08 ->
  // 4.61; 32 columns were slower for every kind (DGER 4.68).
  constexpr int64_t cols = K == K_SYR2 ? 16 : (K == K_SYR || K == K_HER2) ? 8 : 4;
  dim3 grid;
  if (TRI) {
    p.cpc = cols;  // divides RB, as the tile enumeration needs
    const int64_t KT = RB / cols;
    int64_t tiles;
    if (p.lower) {
      tiles = KT * (row_blocks * (row_blocks + 1) / 2);
    } else {
      const int64_t groups = cdiv(cdiv(n, cols), KT);
      tiles = KT * (groups * (groups + 1) / 2);
    }
    if (tiles > 0x7fffffff) return CORAL_ERR_UNSUPPORTED;
    grid = dim3((unsigned)tiles);
  } else {
    p.cpc = std::max<int64_t>(cols, cdiv(n, 65535));
    if (row_blocks > 0x7fffffff) return CORAL_ERR_UNSUPPORTED;
    grid = dim3((unsigned)row_blocks, (unsigned)cdiv(n, p.cpc));
  }
  const int slot = timing_begin(TK_GEMV, s);
  hipLaunchKernelGGL((rank_kernel<E, K>), grid, dim3(kThreads), 0, s, p);
  CORAL_LAUNCH_CHECK();
  timing_end(slot, s);
  return CORAL_OK;
}

constexpr ArgPos kGerPos = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9};   // (m, n, alpha, x, incx, y, incy, a, lda)
constexpr ArgPos kSyrPos = {1, 2, 2, 3, 4, 5, 0, 0, 6, 7};   // (uplo, n, alpha, x, incx, a, lda)
constexpr ArgPos kSyr2Pos = {1, 2, 2, 3, 4, 5, 6, 7, 8, 9};  // (uplo, n, alpha, x, incx, y, incy, a, lda)

}  // namespace rk
}  // namespace coral

using namespace coral;
using namespace coral::rk;

extern "C" {

int coral_sger(int64_t m, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
               float *a, int64_t lda, coral_stream_t stream) {
  return rank_impl<float, K_GER>(kGerPos, 0, m, n, alpha, x, incx, y, incy, a, lda, stream);
}
int coral_dger(int64_t m, int64_t n, double alpha, const double *x, int64_t incx, const double *y, int64_t incy,
               double *a, int64_t lda, coral_stream_t stream) {
  return rank_impl<double, K_GER>(kGerPos, 0, m, n, alpha, x, incx, y, incy, a, lda, stream);
}

#define CORAL_CGER(NAME, R, E2, MK, KIND)                                                                        \
  int coral_##NAME(int64_t m, int64_t n, const R alpha[2], const R *x, int64_t incx, const R *y, int64_t incy,  \
                   R *a, int64_t lda, coral_stream_t stream) {                                                   \
    if (!alpha) return -3;                                                                                       \
    return rank_impl<E2, KIND>(kGerPos, 0, m, n, MK(alpha[0], alpha[1]), (const E2 *)x, incx, (const E2 *)y,    \
                               incy, (E2 *)a, lda, stream);                                                      \
  }
CORAL_CGER(cgeru, float, float2, make_float2, K_GERU)
CORAL_CGER(cgerc, float, float2, make_float2, K_GERC)
CORAL_CGER(zgeru, double, double2, make_double2, K_GERU)
CORAL_CGER(zgerc, double, double2, make_double2, K_GERC)

int coral_ssyr(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda,
               coral_stream_t stream) {
  return rank_impl<float, K_SYR>(kSyrPos, uplo, n, n, alpha, x, incx, nullptr, 1, a, lda, stream);
}
int coral_dsyr(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda,
               coral_stream_t stream) {
  return rank_impl<double, K_SYR>(kSyrPos, uplo, n, n, alpha, x, incx, nullptr, 1, a, lda, stream);
}
int coral_ssyr2(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, const float *y, int64_t incy,
                float *a, int64_t lda, coral_stream_t stream) {
  return rank_impl<float, K_SYR2>(kSyr2Pos, uplo, n, n, alpha, x, incx, y, incy, a, lda, stream);
}
int coral_dsyr2(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream) {
  return rank_impl<double, K_SYR2>(kSyr2Pos, uplo, n, n, alpha, x, incx, y, incy, a, lda, stream);
}
int coral_cher(int32_t uplo, int64_t n, float alpha, const float *x, int64_t incx, float *a, int64_t lda,
               coral_stream_t stream) {
  return rank_impl<float2, K_HER>(kSyrPos, uplo, n, n, make_float2(alpha, 0.0f), (const float2 *)x, incx, nullptr,
                                  1, (float2 *)a, lda, stream);
}
int coral_zher(int32_t uplo, int64_t n, double alpha, const double *x, int64_t incx, double *a, int64_t lda,
               coral_stream_t stream) {
  return rank_impl<double2, K_HER>(kSyrPos, uplo, n, n, make_double2(alpha, 0.0), (const double2 *)x, incx,
                                   nullptr, 1, (double2 *)a, lda, stream);
}
int coral_cher2(int32_t uplo, int64_t n, const float alpha[2], const float *x, int64_t incx, const float *y,
                int64_t incy, float *a, int64_t lda, coral_stream_t stream) {
  if (!alpha) return -3;
  return rank_impl<float2, K_HER2>(kSyr2Pos, uplo, n, n, make_float2(alpha[0], alpha[1]), (const float2 *)x, incx,
                                   (const float2 *)y, incy, (float2 *)a, lda, stream);
}
int coral_zher2(int32_t uplo, int64_t n, const double alpha[2], const double *x, int64_t incx, const double *y,
                int64_t incy, double *a, int64_t lda, coral_stream_t stream) {
  if (!alpha) return -3;
  return rank_impl<double2, K_HER2>(kSyr2Pos, uplo, n, n, make_double2(alpha[0], alpha[1]), (const double2 *)x,
                                    incx, (const double2 *)y, incy, (double2 *)a, lda, stream);
}

}  // extern "C"
