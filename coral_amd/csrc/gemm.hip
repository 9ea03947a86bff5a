// Level-3 GEMM for gfx950: s/d/c/zgemm plus the split-precision BF16x3 forms.
//
// Design (replaces CORAL's Goto/BLIS 5-loop, coral-aarch64/src/level3/sgemm_nn.rs:10-99,
// its packers f32_packers.rs:115-333 and the 8x12 NEON microkernel
// microkernel/f32_mrxnr.rs:99-325):
//
//  1. A pack kernel copies op(A) (m x k) and op(B) (k x n, stored by columns) into
//     "fragment-ordered" tiles.  Transposition, conjugation, complex expansion,
//     power-of-two scaling and the BF16 hi/lo split all happen here, so the MFMA
//     kernel sees a single layout.  Tiles are zero padded, so the main loop has no
//     bounds checks.
//  2. The MFMA kernel streams packed tiles HBM -> registers -> LDS (two stages,
//     one barrier per K tile) and feeds 16x16 MFMA tiles from LDS with one
//     conflict-free ds_read_b128 per lane per fragment.
//  3. The epilogue applies alpha/beta (never reading C when beta == 0, as
//     sgemm_nn.rs:29-45 and f32_mrxnr.rs:164-176 do), row/column scales and, for
//     complex 4M, pairs real/imaginary columns across neighbouring lanes.
//
// Fragment-ordered tile layout.  A tile covers R rows x BK k-values.  One
// "fragment" is 16 rows x KF k-values = 64 lanes x 16 bytes:
//     lane l holds row (l % 16), k-values (l / 16) * VE + t, t < VE,
// where VE = 16 / sizeof(element) and KF = 4 * VE.  A tile is stored as
// [kc < BK/KF][rg < R/16][plane < PL][lane < 64][VE].  The MFMA k index is
// permutation invariant as long as A and B use the same (lane, t) -> k map, so
// every 16-byte vector feeds VE 16x16x4 MFMAs (f32/f64) or one 16x16x32 BF16 MFMA.
#include "coral_common.hpp"
#include "mfma.hpp"

#include <algorithm>
#include <string.h>

namespace coral {
namespace gemm {

// ---------------------------------------------------------------------------
// operand fetchers: value of the packed operand at (r, k) for batch slot z.
// For A, r is a row of op(A); for B, r is a column of op(B).
// ---------------------------------------------------------------------------
template <typename T>
struct FetchReal {  // element p[r*sr + k*sk]
  const T *p;
  int64_t sr, sk;
  __device__ __forceinline__ T operator()(int64_t r, int64_t k, int) const { return p[r * sr + k * sk]; }
};

// Complex 4M expansion of A: A' = [Re(opA) | Im(opA)], an R x 2K real matrix.
template <typename T>
struct FetchCplxA4M {
  const T *p;
  int64_t sr, sk, K;
  bool conj;
  // one load: the element's part is chosen by its address, the sign afterwards
  __device__ __forceinline__ T operator()(int64_t r, int64_t kk, int) const {
    const bool im = kk >= K;
    const int64_t k = im ? kk - K : kk;
    const T v = p[2 * (r * sr + k * sk) + (im ? 1 : 0)];
    return (im && conj) ? -v : v;
  }
};

// Complex 4M expansion of B: column 2j = [Re b_j ; -Im b_j], column 2j+1 = [Im b_j ; Re b_j],
// so [Re(opA) | Im(opA)] * B' interleaves Re(C) and Im(C) column by column.
template <typename T>
struct FetchCplxB4M {
  const T *p;
  int64_t sr, sk, K;
  bool conj;
  // top: part 0 -> Re, part 1 -> Im'; bottom: part 0 -> -Im', part 1 -> Re, with
  // Im' = -Im for a conjugated operand.  One load: the part is (part != bottom),
  // negated for Im' at the bottom and for -Im at the top of a conjugated operand.
  __device__ __forceinline__ T operator()(int64_t jj, int64_t kk, int) const {
    const int64_t j = jj >> 1;
    const bool part = jj & 1;
    const bool bot = kk >= K;
    const int64_t k = bot ? kk - K : kk;
    const bool use_im = part != bot;
    const T v = p[2 * (j * sr + k * sk) + (use_im ? 1 : 0)];
    const bool neg = use_im && (bot != conj);
    return neg ? -v : v;
  }
};

// Complex 3M (Gauss) operand: element (r, k) at p[2 * (r*sr + k*sk)], conjugated
// if conj.  pack_cplx3m_kernel forms its planes Re, Im and Re + Im.
template <typename T>
struct FetchCplx3M {
  const T *p;
  int64_t sr, sk;
  bool conj;
};

// ---------------------------------------------------------------------------
// pack kernels
// ---------------------------------------------------------------------------
// Row scaling for the split form.  mx[r] holds the IEEE bits of max_k |f(r,k)|
// (non-negative floats order like their bit patterns, so atomicMax on the bits is
// an exact, order-independent max; NaN is skipped by fmaxf).  The row exponent is
// e[r] = floor(log2(mx[r])); row r is scaled by 2^(SCALE_EXP - e[r]), so its max
// magnitude lies in [2^SCALE_EXP, 2^(SCALE_EXP+1)); zero or non-finite rows keep e = 0.
__device__ __forceinline__ int exp_of_max(uint32_t bits) {
  const float mx = __uint_as_float(bits);
  return (mx > 0.f && mx <= 3.402823466e38f) ? ilogbf(mx) : 0;
}

__global__ void zero_u32_kernel(uint32_t *p, int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0u;
}

// ---------------------------------------------------------------------------
// Real operands (the hot path).  One "job" per operand: element (r, k) at
// p[r*sr + k*sk] with sr == 1 (rows contiguous: op(A) = A, op(B) = B^T) or
// sk == 1 (k contiguous: op(A) = A^T, op(B) = B).  A and B run in one launch:
// workgroups [0, a.blocks) take job a, the rest job b.
//
// Workgroup blocks are shaped so every global read is a 16-byte load and each
// wave reads whole 1 KiB (rows contiguous) or 512-byte (k contiguous) runs:
//   rows contiguous: RW = 64*LV rows x BK k (one K tile), LV = 16 / sizeof(T)
//   k contiguous:    64 rows x KW = 32*LV k (several K tiles)
// and consecutive workgroups walk the contiguous direction.
// ---------------------------------------------------------------------------
template <typename T>
struct RealJob {
  v4u *dst;                // packed tiles (pack) / unused (absmax)
  const T *p;
  int64_t sr, sk, R, K;
  int64_t KT;              // K tiles of the packed operand
  int64_t Rp, Kp;          // padded extents: cdiv(R, br) * br, KT * bk
  uint32_t *mx;            // split form: max|.| bits per row (absmax or a scanning pack writes, pack reads);
                           // may be null
  int br;                  // tile rows
  int rows_contig;         // sr == 1
  int vec;                 // base and stride allow 16-byte loads
  int64_t rblocks, kblocks;
  int64_t blocks;          // rblocks * kblocks
};

// 16-byte (or guarded scalar) read of LV consecutive elements along the contiguous
// index.  The operand is streamed once per pass, so the loads are nontemporal
// (measured on 2 GiB: 7.1 TB/s vs 6.2 TB/s for ordinary loads).
template <typename T, int LV>
__device__ __forceinline__ void load_run(const RealJob<T> &j, int64_t r, int64_t k, T (&v)[LV]) {
  if (j.rows_contig) {
    if (j.vec && r + LV <= j.R && k < j.K) {
      const v4u x = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(j.p + r + k * j.sk));
      __builtin_memcpy(v, &x, 16);
      return;
    }
#pragma unroll
    for (int e = 0; e < LV; ++e) v[e] = (r + e < j.R && k < j.K) ? j.p[r + e + k * j.sk] : T(0);
  } else {
    if (j.vec && r < j.R && k + LV <= j.K) {
      const v4u x = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(j.p + r * j.sr + k));
      __builtin_memcpy(v, &x, 16);
      return;
    }
#pragma unroll
    for (int e = 0; e < LV; ++e) v[e] = (r < j.R && k + e < j.K) ? j.p[r * j.sr + k + e] : T(0);
  }
}

// The 16-byte nontemporal load of load_run without its guards, for workgroups whose
// whole block lies inside the operand.  A guarded load is a branch, and at each
// branch join the compiler waits for every load issued so far; unguarded, all of a
// thread's loads of the block are in flight at once.
template <typename T, int LV>
__device__ __forceinline__ void load_run_full(const RealJob<T> &j, int64_t r, int64_t k, T (&v)[LV]) {
  const T *src = j.rows_contig ? j.p + r + k * j.sk : j.p + r * j.sr + k;
  const v4u x = __builtin_nontemporal_load(reinterpret_cast<const v4u *>(src));
  __builtin_memcpy(v, &x, 16);
}

// SCAN (split forms whose 16-bit type has FP32's exponent range): the operand is
// split unscaled and the kernel atomicMax-es each row's max |x| into mx, from the
// block it holds in LDS.  gate (may be null): the workgroup packs only if *gate is
// non-zero, so a re-pack can be queued behind a device-side decision.
template <class TR, int BK, bool SCAN>
__global__ __launch_bounds__(256) void pack_real_kernel(const RealJob<typename TR::T> ja,
                                                        const RealJob<typename TR::T> jb, const uint32_t *gate) {
  using T = typename TR::T;
  static_assert(!SCAN || TR::SPLIT, "only split forms scan for row maxima");
  if (gate && *gate == 0u) return;
  constexpr int LV = 16 / sizeof(T);            // elements per 16-byte load
  constexpr int VE = TR::VE, PL = TR::PL, KF = TR::KF;
  constexpr int SV = TR::SPLIT ? VE : LV;       // source elements per fragment vector
  constexpr int RW = 64 * LV, KW = 32 * LV;
  static_assert(KW % BK == 0, "k-contiguous block must hold whole K tiles");
  // LDS rows padded so the fragment gather (lanes 16 apart read k values SV apart)
  // hits different banks
  constexpr int PADR = SV == 8 ? 2 : (SV == 4 ? 4 : 8);
  constexpr int ROWS_LDS = BK * (RW + PADR), KC_LDS = 64 * (KW + LV);
  __shared__ __attribute__((aligned(16))) T s[ROWS_LDS > KC_LDS ? ROWS_LDS : KC_LDS];

  const bool first = blockIdx.x < ja.blocks;
  const RealJob<T> &j = first ? ja : jb;
  const int64_t blk = first ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - ja.blocks;
  const int t = threadIdx.x;
  int64_t r0, k0;
  int rspan, kspan, ld;
  if (j.rows_contig) {
    r0 = (blk % j.rblocks) * RW;
    k0 = (blk / j.rblocks) * BK;
    rspan = RW; kspan = BK; ld = RW + PADR;
    T v[BK / 4][LV];
    if (j.vec && r0 + RW <= j.R && k0 + BK <= j.K) {  // whole block inside: unguarded loads
#pragma unroll
      for (int c = 0; c < BK / 4; ++c) load_run_full<T, LV>(j, r0 + LV * (t & 63), k0 + (t >> 6) + 4 * c, v[c]);
    } else {
#pragma unroll
      for (int c = 0; c < BK / 4; ++c) load_run<T, LV>(j, r0 + LV * (t & 63), k0 + (t >> 6) + 4 * c, v[c]);
    }
#pragma unroll
    for (int c = 0; c < BK / 4; ++c) {
      const int kl = (t >> 6) + 4 * c, rl = LV * (t & 63);
#pragma unroll
      for (int e = 0; e < LV; ++e) s[kl * ld + rl + e] = v[c][e];  // s[k][r]
    }
  } else {
    k0 = (blk % j.kblocks) * KW;
    r0 = (blk / j.kblocks) * 64;
    rspan = 64; kspan = KW; ld = KW + LV;
    T v[8][LV];
    if (j.vec && r0 + 64 <= j.R && k0 + KW <= j.K) {  // whole block inside: unguarded loads
#pragma unroll
      for (int c = 0; c < 8; ++c) load_run_full<T, LV>(j, r0 + (t >> 5) + 8 * c, k0 + LV * (t & 31), v[c]);
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) load_run<T, LV>(j, r0 + (t >> 5) + 8 * c, k0 + LV * (t & 31), v[c]);
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int rl = (t >> 5) + 8 * c, kl = LV * (t & 31);
      v4u x;
      __builtin_memcpy(&x, v[c], 16);
      *reinterpret_cast<v4u *>(&s[rl * ld + kl]) = x;  // s[r][k], 16-byte aligned rows
    }
  }
  __syncthreads();

  if constexpr (SCAN) {  // row maxima of the block (zeros pad it beyond R and K)
    if (j.rows_contig) {   // s[k][r]: thread t scans row t
      float m = 0.f;
#pragma unroll 8
      for (int kl = 0; kl < BK; ++kl) m = fmaxf(m, fabsf((float)s[kl * ld + t]));
      if (r0 + t < j.R) atomicMax(&j.mx[r0 + t], __float_as_uint(m));
    } else {               // s[r][k]: four threads per row, k interleaved, then combined
      const int rl = t >> 2, kq = t & 3;
      float m = 0.f;
#pragma unroll 8
      for (int i = 0; i < KW / 4; ++i) m = fmaxf(m, fabsf((float)s[rl * ld + kq + 4 * i]));
      m = fmaxf(m, __shfl_xor(m, 1, 64));
      m = fmaxf(m, __shfl_xor(m, 2, 64));
      if (kq == 0 && r0 + rl < j.R) atomicMax(&j.mx[r0 + rl], __float_as_uint(m));
    }
  }

  const int RG = j.br / 16;
  const int units = (rspan / 16) * (kspan / KF) * 64;
  for (int u = t; u < units; u += 256) {
    const int lane = u & 63, rest = u >> 6;
    const int rgl = rest % (rspan / 16), kcl = rest / (rspan / 16);
    const int rl = rgl * 16 + (lane & 15), kl = kcl * KF + (lane >> 4) * SV;
    const int64_t gr = r0 + rl, gk = k0 + kl;
    if (gr >= j.Rp || gk >= j.Kp) continue;
    T v[SV];
    if (j.rows_contig) {
#pragma unroll
      for (int e = 0; e < SV; ++e) v[e] = s[(kl + e) * ld + rl];
    } else {
#pragma unroll
      for (int e = 0; e < SV; e += LV) {
        const v4u x = *reinterpret_cast<const v4u *>(&s[rl * ld + kl + e]);
        __builtin_memcpy(&v[e], &x, 16);
      }
    }
    const int64_t tile = (gr / j.br) * j.KT + gk / BK;
    const int rg = (int)(gr % j.br) / 16, kc = (int)(gk % BK) / KF;
    v4u *out = j.dst + tile * ((int64_t)(BK / KF) * RG * 64 * PL) + ((int64_t)(kc * RG + rg) * PL) * 64 +
               ((gr & 15) + 16 * ((gk % KF) / SV));
    if constexpr (!TR::SPLIT) {
      out[0] = __builtin_bit_cast(v4u, v);
    } else {
      const int sc = (!SCAN && gr < j.R && j.mx) ? TR::SCALE_EXP - exp_of_max(j.mx[gr]) : 0;
      uint32_t hi[VE], lo[VE];
#pragma unroll
      for (int e = 0; e < VE; ++e) TR::split(ldexpf((float)v[e], sc), hi[e], lo[e]);
      v4u vh, vl;
      vh.x = hi[0] | (hi[1] << 16); vh.y = hi[2] | (hi[3] << 16);
      vh.z = hi[4] | (hi[5] << 16); vh.w = hi[6] | (hi[7] << 16);
      vl.x = lo[0] | (lo[1] << 16); vl.y = lo[2] | (lo[3] << 16);
      vl.z = lo[4] | (lo[5] << 16); vl.w = lo[6] | (lo[7] << 16);
      out[0] = vh;
      out[64] = vl;
    }
  }
}

// Complex 3M operands for the split forms: plane z = 0 holds Re, z = 1 Im and
// z = 2 Re + Im of op(X) (Im negated for a conjugated operand), each packed like a
// real operand.  A workgroup stages one block of complex elements in LDS with
// 16-byte loads (two elements each) and writes the block's tiles of all three
// planes, so the operand is read once: rows contiguous, 128 rows x BK k; k
// contiguous, 64 rows x 64 k.  SCAN and gate as in pack_real_kernel, with the
// three planes' row maxima at mx[z * mx_zstride + r]; !PACK: the maxima only.
struct Cplx3MJob {
  v4u *dst;
  int64_t dst_zstride;     // v4u per plane
  const float *p;          // interleaved (re, im)
  int64_t sr, sk, R, K;    // element (r, k) at p[2 * (r*sr + k*sk)]
  int64_t KT, Rp, Kp;
  uint32_t *mx;
  int64_t mx_zstride;
  int br, rows_contig, vec, conj;
  int64_t rblocks, kblocks, blocks;
};

template <class TR, int BK, bool SCAN, bool PACK>
__global__ __launch_bounds__(256) void pack_cplx3m_kernel(const Cplx3MJob ja, const Cplx3MJob jb,
                                                          const uint32_t *gate) {
  static_assert(TR::SPLIT, "3M planes are packed for the split forms");
  static_assert(SCAN || PACK, "nothing to do");
  constexpr int VE = TR::VE, PL = TR::PL, KF = TR::KF;
  constexpr int RW = 128, KW = 64, PAD = 1;
  static_assert(KW % BK == 0, "k-contiguous block must hold whole K tiles");
  constexpr int ROWS_LDS = BK * (RW + PAD), KC_LDS = 64 * (KW + PAD);
  __shared__ float2 s[ROWS_LDS > KC_LDS ? ROWS_LDS : KC_LDS];
  if (gate && *gate == 0u) return;

  const bool first = blockIdx.x < ja.blocks;
  const Cplx3MJob &j = first ? ja : jb;
  const int64_t blk = first ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - ja.blocks;
  const int t = threadIdx.x;
  int64_t r0, k0;
  int rspan, kspan, ld;
  if (j.rows_contig) {
    r0 = (blk % j.rblocks) * RW;
    k0 = (blk / j.rblocks) * BK;
    rspan = RW; kspan = BK; ld = RW + PAD;
    const int rl = 2 * (t & 63);
    float2 v[BK / 4][2];
    if (j.vec && r0 + RW <= j.R && k0 + BK <= j.K) {  // whole block inside: unguarded loads
#pragma unroll
      for (int c = 0; c < BK / 4; ++c) {
        const v4u x = __builtin_nontemporal_load(
            reinterpret_cast<const v4u *>(j.p + 2 * (r0 + rl + (k0 + (t >> 6) + 4 * c) * j.sk)));
        __builtin_memcpy(&v[c][0], &x, 16);
      }
    } else {
#pragma unroll
      for (int c = 0; c < BK / 4; ++c) {
        const int64_t k = k0 + (t >> 6) + 4 * c;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t r = r0 + rl + e;
          v[c][e] = (r < j.R && k < j.K) ? make_float2(j.p[2 * (r + k * j.sk)], j.p[2 * (r + k * j.sk) + 1])
                                         : make_float2(0.f, 0.f);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < BK / 4; ++c) {
      const int kl = (t >> 6) + 4 * c;
      s[kl * ld + rl] = v[c][0];
      s[kl * ld + rl + 1] = v[c][1];
    }
  } else {
    k0 = (blk % j.kblocks) * KW;
    r0 = (blk / j.kblocks) * 64;
    rspan = 64; kspan = KW; ld = KW + PAD;
    const int kl = 2 * (t & 31);
    float2 v[8][2];
    if (j.vec && r0 + 64 <= j.R && k0 + KW <= j.K) {  // whole block inside: unguarded loads
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const v4u x = __builtin_nontemporal_load(
            reinterpret_cast<const v4u *>(j.p + 2 * ((r0 + (t >> 5) + 8 * c) * j.sr + k0 + kl)));
        __builtin_memcpy(&v[c][0], &x, 16);
      }
    } else {
#pragma unroll
      for (int c = 0; c < 8; ++c) {
        const int64_t r = r0 + (t >> 5) + 8 * c;
#pragma unroll
        for (int e = 0; e < 2; ++e) {
          const int64_t k = k0 + kl + e;
          v[c][e] = (r < j.R && k < j.K) ? make_float2(j.p[2 * (r * j.sr + k)], j.p[2 * (r * j.sr + k) + 1])
                                         : make_float2(0.f, 0.f);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 8; ++c) {
      const int rl = (t >> 5) + 8 * c;
      s[rl * ld + kl] = v[c][0];
      s[rl * ld + kl + 1] = v[c][1];
    }
  }
  __syncthreads();

  const float isg = j.conj ? -1.f : 1.f;
  if constexpr (SCAN) {  // the three planes' row maxima (zeros pad the block)
    float m0 = 0.f, m1 = 0.f, m2 = 0.f;
    int64_t r = -1;
    if (j.rows_contig) {
      if (t < RW) {
        for (int kl = 0; kl < BK; ++kl) {
          const float2 x = s[kl * ld + t];
          const float re = x.x, im = isg * x.y;
          m0 = fmaxf(m0, fabsf(re));
          m1 = fmaxf(m1, fabsf(im));
          m2 = fmaxf(m2, fabsf(re + im));
        }
        r = r0 + t;
      }
    } else {
      const int rl = t >> 2, kq = t & 3;
#pragma unroll 4
      for (int i = 0; i < KW / 4; ++i) {
        const float2 x = s[rl * ld + kq + 4 * i];
        const float re = x.x, im = isg * x.y;
        m0 = fmaxf(m0, fabsf(re));
        m1 = fmaxf(m1, fabsf(im));
        m2 = fmaxf(m2, fabsf(re + im));
      }
#pragma unroll
      for (int o = 1; o <= 2; o <<= 1) {
        m0 = fmaxf(m0, __shfl_xor(m0, o, 64));
        m1 = fmaxf(m1, __shfl_xor(m1, o, 64));
        m2 = fmaxf(m2, __shfl_xor(m2, o, 64));
      }
      if (kq == 0) r = r0 + rl;
    }
    if (r >= 0 && r < j.R) {
      atomicMax(&j.mx[r], __float_as_uint(m0));
      atomicMax(&j.mx[j.mx_zstride + r], __float_as_uint(m1));
      atomicMax(&j.mx[2 * j.mx_zstride + r], __float_as_uint(m2));
    }
  }
  if constexpr (!PACK) return;

  const int RG = j.br / 16;
  const int units = (rspan / 16) * (kspan / KF) * 64;
  for (int u = t; u < units; u += 256) {
    const int lane = u & 63, rest = u >> 6;
    const int rgl = rest % (rspan / 16), kcl = rest / (rspan / 16);
    const int rl = rgl * 16 + (lane & 15), kl = kcl * KF + (lane >> 4) * VE;
    const int64_t gr = r0 + rl, gk = k0 + kl;
    if (gr >= j.Rp || gk >= j.Kp) continue;
    float re[VE], im[VE];
#pragma unroll
    for (int e = 0; e < VE; ++e) {
      const float2 x = j.rows_contig ? s[(kl + e) * ld + rl] : s[rl * ld + kl + e];
      re[e] = x.x;
      im[e] = isg * x.y;
    }
    const int64_t tile = (gr / j.br) * j.KT + gk / BK;
    const int rg = (int)(gr % j.br) / 16, kc = (int)(gk % BK) / KF;
    v4u *out = j.dst + tile * ((int64_t)(BK / KF) * RG * 64 * PL) + ((int64_t)(kc * RG + rg) * PL) * 64 +
               ((gr & 15) + 16 * ((gk % KF) / VE));
#pragma unroll
    for (int z = 0; z < 3; ++z) {
      const int sc = (!SCAN && gr < j.R && j.mx) ? TR::SCALE_EXP - exp_of_max(j.mx[z * j.mx_zstride + gr]) : 0;
      uint32_t hi[VE], lo[VE];
#pragma unroll
      for (int e = 0; e < VE; ++e) {
        const float x = z == 0 ? re[e] : (z == 1 ? im[e] : re[e] + im[e]);
        TR::split(ldexpf(x, sc), hi[e], lo[e]);
      }
      v4u vh, vl;
      vh.x = hi[0] | (hi[1] << 16); vh.y = hi[2] | (hi[3] << 16);
      vh.z = hi[4] | (hi[5] << 16); vh.w = hi[6] | (hi[7] << 16);
      vl.x = lo[0] | (lo[1] << 16); vl.y = lo[2] | (lo[3] << 16);
      vl.z = lo[4] | (lo[5] << 16); vl.w = lo[6] | (lo[7] << 16);
      out[z * j.dst_zstride] = vh;
      out[z * j.dst_zstride + 64] = vl;
    }
  }
}

// max_k |x(r, k)| per row for the split form (FP32 operands), both jobs in one
// launch; mx must hold zeros.  Rows contiguous: a workgroup covers 256 rows x 256 k
// (lanes take 4 rows each with 16-byte loads, the 4 waves split k), then one
// atomicMax per row.  k contiguous: one wave per row and 8192-k chunk.  Up to 16
// loads per thread in flight.  At 8192^3 this reads A and B in 86 us, against 90
// with half the chunk on both sides and 8 loads in flight.
__global__ __launch_bounds__(256) void absmax_real_kernel(const RealJob<float> ja, const RealJob<float> jb) {
  constexpr int KCH_ROWS = 256, KCH_K = 8192;
  __shared__ float red[4][256];
  const bool first = blockIdx.x < ja.blocks;
  const RealJob<float> &j = first ? ja : jb;
  const int64_t blk = first ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - ja.blocks;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  if (j.rows_contig) {
    const int64_t r0 = (blk % j.rblocks) * 256, k0 = (blk / j.rblocks) * KCH_ROWS;
    const int64_t r = r0 + 4 * lane;
    float m[4] = {0.f, 0.f, 0.f, 0.f};
    if (j.vec && r0 + 256 <= j.R && k0 + KCH_ROWS <= j.K) {  // whole block inside: unguarded loads
#pragma unroll 16
      for (int c = 0; c < KCH_ROWS / 4; ++c) {
        float v[4];
        load_run_full<float, 4>(j, r, k0 + w + 4 * c, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) m[e] = fmaxf(m[e], fabsf(v[e]));
      }
    } else {
#pragma unroll 16
      for (int c = 0; c < KCH_ROWS / 4; ++c) {
        float v[4];
        load_run<float, 4>(j, r, k0 + w + 4 * c, v);
#pragma unroll
        for (int e = 0; e < 4; ++e) m[e] = fmaxf(m[e], fabsf(v[e]));
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) red[w][4 * lane + e] = m[e];
    __syncthreads();
    const float mm = fmaxf(fmaxf(red[0][t], red[1][t]), fmaxf(red[2][t], red[3][t]));
    if (r0 + t < j.R) atomicMax(&j.mx[r0 + t], __float_as_uint(mm));
  } else {
    const int64_t r = (blk % j.rblocks) * 4 + w, k0 = (blk / j.rblocks) * KCH_K;
    float m = 0.f;
    if (j.vec && r < j.R && k0 + KCH_K <= j.K) {  // this wave's whole run inside: unguarded loads
#pragma unroll 16
      for (int c = 0; c < KCH_K / 256; ++c) {
        float v[4];
        load_run_full<float, 4>(j, r, k0 + 256 * c + 4 * lane, v);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
      }
    } else if (r < j.R) {
#pragma unroll 4
      for (int c = 0; c < KCH_K / 256; ++c) {
        float v[4];
        load_run<float, 4>(j, r, k0 + 256 * c + 4 * lane, v);
        m = fmaxf(m, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
      }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off, 64));
    if (lane == 0 && r < j.R) atomicMax(&j.mx[r], __float_as_uint(m));
  }
}

// After a scanning pack: *gate = 1 if some row maximum in mxa or mxb lies
// outside [2^-40, 2^40] and is not zero (Inf included), else 0.  Inside that range
// the unscaled split is the scaled one times a power of two: hi, lo, every product
// (>= 2^-80 * 2^-18 against 2^-126) and every partial sum (<= 2^31 * 2^80 against
// 2^128) stay normal and finite, so power-of-two scaling commutes with each
// rounding and changes no result bit.  One thread per maximum; *gate must hold 0.
__global__ __launch_bounds__(1024) void split_gate_kernel(const uint32_t *__restrict__ mxa, int64_t na,
                                                          const uint32_t *__restrict__ mxb, int64_t nb,
                                                          uint32_t *__restrict__ gate) {
  constexpr uint32_t LO = 0x2b800000u /* 2^-40 */, HI = 0x53800000u /* 2^40 */;
  const int64_t i = (int64_t)blockIdx.x * 1024 + threadIdx.x;
  int bad = 0;
  if (i < na + nb) {
    const uint32_t b = i < na ? mxa[i] : mxb[i - na];
    bad = b != 0u && (b < LO || b > HI);
  }
  if (__syncthreads_or(bad) && threadIdx.x == 0) atomicOr(gate, 1u);
}

// One thread writes the PL vectors (16 bytes each) of one (tile, kc, rg, lane) unit
// of a plain (unsplit) operand; used for the complex 4M expansions.
template <class TR, int BR, int BK, class F>
__global__ void pack_kernel(v4u *__restrict__ dst, int64_t R, int64_t K, int64_t KT, int64_t units, F f) {
  static_assert(!TR::SPLIT, "split operands are packed by pack_real_kernel / pack_cplx3m_kernel");
  constexpr int VE = TR::VE, PL = TR::PL, KF = TR::KF, RG = BR / 16, KC = BK / KF;
  constexpr int UNITS_PER_TILE = KC * RG * 64;
  int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= units) return;
  int64_t tile = u / UNITS_PER_TILE;
  int w = (int)(u % UNITS_PER_TILE);
  int lane = w & 63;
  int rest = w >> 6;
  int rg = rest % RG;
  int kc = rest / RG;
  int64_t tr = tile / KT, tk = tile % KT;
  int64_t r = tr * BR + rg * 16 + (lane & 15);
  int64_t k0 = tk * BK + kc * KF + (lane >> 4) * VE;
  v4u *out = dst + tile * (int64_t)(UNITS_PER_TILE * PL) + ((int64_t)(kc * RG + rg) * PL) * 64 + lane;
  using T = typename TR::T;
  // loads from indices clamped into the operand, all issued before any is used;
  // elements outside it are zeroed afterwards (a guarded load per element is a
  // branch, and each branch join waits for every load issued before it)
  const int64_t rc = r < R ? r : R - 1;
  T v[VE];
#pragma unroll
  for (int t = 0; t < VE; ++t) v[t] = f(rc, k0 + t < K ? k0 + t : K - 1, 0);
#pragma unroll
  for (int t = 0; t < VE; ++t)
    if (r >= R || k0 + t >= K) v[t] = T(0);
  out[0] = __builtin_bit_cast(v4u, v);
}

// ---------------------------------------------------------------------------
// MFMA kernel
// ---------------------------------------------------------------------------
enum Epi { EPI_REAL = 0, EPI_CPLX = 1, EPI_STORE_Z = 2 };

struct GemmArgs {
  const v4u *Ap, *Bp;
  int64_t a_zstride, b_zstride;  // in 16-byte vectors
  int64_t M, N;                  // output extent of the real product (N = 2n for complex 4M)
  int64_t KT;                    // number of BK tiles
  int tiles_m, tiles_n;
  void *C;
  int64_t ldc, c_zstride;
  double ar, ai, br, bi;
  int beta_zero;
  const uint32_t *row_exp, *col_exp;  // split form only: max|.| bits per row of op(A) / column of op(B)
  int64_t re_zstride, ce_zstride;
  const uint32_t *gate;  // split form: if set and *gate == 0, the operands were packed unscaled
};

// tile id -> (tm, tn): spread consecutive ids over the 8 XCDs so neighbours land
// on one XCD, then walk column groups of 8 tile rows (L2 reuse of B panels).
__device__ __forceinline__ void map_tile(int bid, int tiles_m, int tiles_n, int &tm, int &tn) {
  const int total = tiles_m * tiles_n;
  int w = bid;
  if ((total & 7) == 0) w = (w & 7) * (total >> 3) + (w >> 3);
  constexpr int GROUP = 8;
  const int per_group = GROUP * tiles_n;
  const int grp = w / per_group;
  const int first_m = grp * GROUP;
  const int gsz = min(tiles_m - first_m, GROUP);
  const int win = w - grp * per_group;
  tm = first_m + win % gsz;
  tn = win / gsz;
}

// Writes a wave's FM x FN accumulator blocks; row_base/col_base is the wave's
// first output row/column (in the real product's index space).
template <class TR, int EPI, int FM, int FN>
__device__ __forceinline__ void epilogue(const typename TR::acc_t (&acc)[FM][FN], const GemmArgs &g, int64_t row_base,
                                         int64_t col_base, int lane, int z) {
  using T = typename TR::T;
  const T ar = (T)g.ar, ai = (T)g.ai, br = (T)g.br, bi = (T)g.bi;
  T *C = reinterpret_cast<T *>(g.C);
  const bool scaled = TR::SPLIT && (!g.gate || *g.gate != 0u);
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int64_t col = col_base + j * 16 + (lane & 15);
      // 16x16 accumulator: lane l holds column l%16; register r holds row 4*(l/16)+r
      // (f32/bf16/f16 MFMA) or row l/16+4*r (f64 MFMA).
      const int64_t row0 = row_base + i * 16 + (TR::ROWS_INTERLEAVED ? (lane >> 4) : 4 * (lane >> 4));
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = row0 + (TR::ROWS_INTERLEAVED ? 4 * r : r);
        T v = acc[i][j][r];
        if constexpr (EPI == EPI_CPLX) {
          T pv = __shfl_xor(v, 1, 64);
          const bool odd = lane & 1;
          const T re = odd ? pv : v, im = odd ? v : pv;
          const int64_t cj = col >> 1;
          if (row < g.M && col < g.N) {
            T *cp = C + 2 * (row + cj * g.ldc);
            T ore = ar * re - ai * im, oim = ar * im + ai * re;
            if (!g.beta_zero) {
              T cr = cp[0], ci = cp[1];
              ore += br * cr - bi * ci;
              oim += br * ci + bi * cr;
            }
            cp[odd ? 1 : 0] = odd ? oim : ore;
          }
        } else {
          if (row < g.M && col < g.N) {
            if constexpr (TR::SPLIT) {
              if (scaled) {
                int sh = 0;
                if (g.row_exp) sh += exp_of_max(g.row_exp[z * g.re_zstride + row]) - TR::SCALE_EXP;
                if (g.col_exp) sh += exp_of_max(g.col_exp[z * g.ce_zstride + col]) - TR::SCALE_EXP;
                v = ldexpf(v, sh);
              }
            }
            if constexpr (EPI == EPI_STORE_Z) {
              C[z * g.c_zstride + row + col * g.ldc] = v;
            } else {
              T *cp = C + row + col * g.ldc;
              T o = ar * v;
              if (!g.beta_zero) o += br * (*cp);
              *cp = o;
            }
          }
        }
      }
    }
  }
}

// Register-staged kernel: 4 waves in a 2 x 2 arrangement, packed tiles travel
// HBM -> registers -> LDS.
template <class TR, int BM, int BN, int BK, int EPI>
__global__ __launch_bounds__(256) void gemm_kernel(GemmArgs g) {
  using acc_t = typename TR::acc_t;
  constexpr int PL = TR::PL, KF = TR::KF;
  constexpr int KC = BK / KF;
  constexpr int RGA = BM / 16, RGB = BN / 16;
  constexpr int ATV = KC * RGA * PL * 64;  // vectors per A tile
  constexpr int BTV = KC * RGB * PL * 64;
  constexpr int LA = ATV / 256, LB = BTV / 256;
  static_assert(ATV % 256 == 0 && BTV % 256 == 0, "tile must be a multiple of 4 KiB");
  constexpr int WM = BM / 2, WN = BN / 2, FM = WM / 16, FN = WN / 16;

  extern __shared__ v4u lds[];  // [2][ATV + BTV]

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int z = blockIdx.z;
  int tm, tn;
  map_tile(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);

  const v4u *Ag = g.Ap + z * g.a_zstride + (int64_t)tm * g.KT * ATV + tid;
  const v4u *Bg = g.Bp + z * g.b_zstride + (int64_t)tn * g.KT * BTV + tid;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  // Pipeline: LDS holds tiles kt (read) and kt+1 (written at the end of iteration
  // kt); registers hold tiles in flight from HBM, two tiles ahead, in two slots:
  // slot (t & 1) carries tile t.  Iteration kt computes from LDS, stores tile kt+1
  // (slot (kt+1)&1, loaded two iterations earlier) into the other LDS stage, then
  // reuses that slot for tile kt+3.  The loop is unrolled by two so slot indices
  // stay compile-time constants (register arrays, no scratch).
  const int64_t KT = g.KT;
  v4u ra0[LA], rb0[LB], ra1[LA], rb1[LB];
  auto load_tile = [&](int64_t t, v4u(&ra)[LA], v4u(&rb)[LB]) {
    const v4u *an = Ag + t * ATV;
    const v4u *bn = Bg + t * BTV;
#pragma unroll
    for (int i = 0; i < LA; ++i) ra[i] = an[i * 256];
#pragma unroll
    for (int i = 0; i < LB; ++i) rb[i] = bn[i * 256];
  };
  auto store_tile = [&](int stage, const v4u(&ra)[LA], const v4u(&rb)[LB]) {
    v4u *dst = lds + stage * (ATV + BTV);
#pragma unroll
    for (int i = 0; i < LA; ++i) dst[tid + i * 256] = ra[i];
#pragma unroll
    for (int i = 0; i < LB; ++i) dst[ATV + tid + i * 256] = rb[i];
  };
  auto compute = [&](int stage) {
    const v4u *la = lds + stage * (ATV + BTV);
    const v4u *lb = la + ATV;
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      v4u fa[FM][PL], fb[FN][PL];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int p = 0; p < PL; ++p) fa[i][p] = la[((kc * RGA + wm * FM + i) * PL + p) * 64 + lane];
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int p = 0; p < PL; ++p) fb[j][p] = lb[((kc * RGB + wn * FN + j) * PL + p) * 64 + lane];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) TR::mma(acc[i][j], fa[i], fb[j]);
    }
  };
  // iteration kt: slot s = (kt + 1) & 1 carries tile kt + 1
  auto step = [&](int64_t kt, v4u(&ra)[LA], v4u(&rb)[LB]) {
    compute((int)(kt & 1));
    if (kt + 1 < KT) store_tile((int)((kt + 1) & 1), ra, rb);
    if (kt + 3 < KT) load_tile(kt + 3, ra, rb);
    __syncthreads();
  };

  load_tile(0, ra0, rb0);
  store_tile(0, ra0, rb0);
  if (KT > 1) load_tile(1, ra1, rb1);
  if (KT > 2) load_tile(2, ra0, rb0);
  __syncthreads();
  for (int64_t kt = 0; kt < KT; kt += 2) {
    step(kt, ra1, rb1);                    // stores tile kt+1 (slot 1), refills slot 1 with kt+3
    if (kt + 1 < KT) step(kt + 1, ra0, rb0);  // stores tile kt+2 (slot 0), refills slot 0 with kt+4
  }
  epilogue<TR, EPI, FM, FN>(acc, g, (int64_t)tm * BM + wm * WM, (int64_t)tn * BN + wn * WN, lane, z);
}

#define CORAL_LDS_PTR(p) ((__attribute__((address_space(3))) void *)(p))

// Direct-to-LDS kernel: WVM x WVN waves; packed tiles go HBM -> LDS with
// buffer_load_dwordx4 ... lds (lane l of a wave writes 16 bytes at the wave's base +
// 16 l, which is exactly the fragment order of the packed tile), so no registers
// carry tiles in flight.  Two LDS stages in two separate arrays (the compiler can
// then tell a ds_read of one stage from the DMA into the other and does not wait
// on the DMA); one barrier per K tile: iteration kt issues the DMA of tile kt+1,
// computes tile kt, then waits for its own DMA and meets the other waves.
//
// SCH selects the instruction schedule of one K tile: 0 leaves it to the compiler;
// 1 issues every fragment read of the tile before the first MFMA (a scheduling
// fence keeps them there), so MFMAs start as soon as their own fragments land
// while later reads are still in flight; 2 pipelines the reads row by row.
template <class TR, int BM, int BN, int BK, int WVM, int WVN, int SCH, int EPI>
__global__ __launch_bounds__(64 * WVM * WVN) void gemm_dma_kernel(GemmArgs g) {
  using acc_t = typename TR::acc_t;
  constexpr int NT = 64 * WVM * WVN;
  constexpr int PL = TR::PL, KF = TR::KF;
  constexpr int KC = BK / KF;
  constexpr int RGA = BM / 16, RGB = BN / 16;
  constexpr int ATV = KC * RGA * PL * 64;
  constexpr int BTV = KC * RGB * PL * 64;
  constexpr int LA = ATV / NT, LB = BTV / NT;
  static_assert(ATV % NT == 0 && BTV % NT == 0, "tile must split evenly over the workgroup");
  constexpr int WM = BM / WVM, WN = BN / WVN, FM = WM / 16, FN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0, "wave tile must be a multiple of 16");

  __shared__ v4u st0[ATV + BTV];
  __shared__ v4u st1[ATV + BTV];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave % WVM, wn = wave / WVM;
  const int z = blockIdx.z;
  int tm, tn;
  map_tile(blockIdx.x, g.tiles_m, g.tiles_n, tm, tn);

  // wave-uniform bases (scalar registers); each lane adds its own 16-byte slot, so
  // the DMA uses the saddr + voffset form and costs no 64-bit VGPR address pairs
  const v4u *Ag = g.Ap + z * g.a_zstride + (int64_t)tm * g.KT * ATV;
  const v4u *Bg = g.Bp + z * g.b_zstride + (int64_t)tn * g.KT * BTV;
  const int wbase = wave * 64;  // this wave's first vector within each NT-vector row

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  // Buffer-to-LDS DMA through a descriptor over the packed tile.  Not global_load_lds:
  // the compiler books that FLAT-encoded form as a possible LDS access on the LGKM
  // counter, which makes it treat that counter as out of order and wait for every
  // outstanding fragment read (lgkmcnt(0)) before each MFMA group; with buffer loads
  // it waits for exactly the reads each MFMA needs.
  auto dma = [&](v4u *st, int64_t t) {
    const __amdgpu_buffer_rsrc_t ra =
        __builtin_amdgcn_make_buffer_rsrc((void *)(Ag + t * ATV), (short)0, ATV * 16, 0x00020000);
    const __amdgpu_buffer_rsrc_t rb =
        __builtin_amdgcn_make_buffer_rsrc((void *)(Bg + t * BTV), (short)0, BTV * 16, 0x00020000);
    const int32_t lane_off = tid * 16;
#pragma unroll
    for (int i = 0; i < LA; ++i)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(ra, CORAL_LDS_PTR(st + i * NT + wbase), 16, lane_off + i * NT * 16, 0,
                                               0, 0);
#pragma unroll
    for (int i = 0; i < LB; ++i)
      __builtin_amdgcn_raw_ptr_buffer_load_lds(rb, CORAL_LDS_PTR(st + ATV + i * NT + wbase), 16,
                                               lane_off + i * NT * 16, 0, 0, 0);
  };
  auto compute = [&](const v4u *st) {
    const v4u *la = st;
    const v4u *lb = st + ATV;
    if constexpr (SCH == 2) {
      // software pipeline over the rows of the wave tile: the fragment of row i+1 is
      // read while the MFMAs of row i run; the group barriers pin that order
      // (0x100 = LDS reads, 0x008 = MFMA).
#pragma unroll
      for (int kc = 0; kc < KC; ++kc) {
        v4u fb[FN][PL], fa[2][PL];
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int p = 0; p < PL; ++p) fb[j][p] = lb[((kc * RGB + wn * FN + j) * PL + p) * 64 + lane];
#pragma unroll
        for (int p = 0; p < PL; ++p) fa[0][p] = la[((kc * RGA + wm * FM) * PL + p) * 64 + lane];
        __builtin_amdgcn_sched_group_barrier(0x100, FN * PL + PL, 0);
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          if (i + 1 < FM) {
#pragma unroll
            for (int p = 0; p < PL; ++p)
              fa[(i + 1) & 1][p] = la[((kc * RGA + wm * FM + i + 1) * PL + p) * 64 + lane];
            __builtin_amdgcn_sched_group_barrier(0x100, PL, 0);
          }
#pragma unroll
          for (int j = 0; j < FN; ++j) TR::mma(acc[i][j], fa[i & 1], fb[j]);
          __builtin_amdgcn_sched_group_barrier(0x008, FN * TR::NMMA, 0);
        }
      }
      return;
    }
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      v4u fa[FM][PL], fb[FN][PL];
#pragma unroll
      for (int j = 0; j < FN; ++j)
#pragma unroll
        for (int p = 0; p < PL; ++p) fb[j][p] = lb[((kc * RGB + wn * FN + j) * PL + p) * 64 + lane];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int p = 0; p < PL; ++p) fa[i][p] = la[((kc * RGA + wm * FM + i) * PL + p) * 64 + lane];
      if constexpr (SCH == 1) __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) TR::mma(acc[i][j], fa[i], fb[j]);
    }
  };
  auto wait_and_meet = [&]() {
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
  };

  const int64_t KT = g.KT;
  dma(st0, 0);
  wait_and_meet();
  for (int64_t kt = 0; kt < KT; kt += 2) {
    if (kt + 1 < KT) dma(st1, kt + 1);
    compute(st0);
    wait_and_meet();
    if (kt + 1 < KT) {
      if (kt + 2 < KT) dma(st0, kt + 2);
      compute(st1);
      wait_and_meet();
    }
  }
  epilogue<TR, EPI, FM, FN>(acc, g, (int64_t)tm * BM + wm * WM, (int64_t)tn * BN + wn * WN, lane, z);
}

// Small real problems (fewer 64x64 output tiles than CUs): one workgroup per 32x32
// tile of C reads op(A) and op(B) straight from the caller's arrays, writes them to
// LDS in fragment order and runs the MFMAs; four wavefronts each own a 16x16 block.
// No pack pass: one launch.  The next K step (BK deep) is loaded into registers while
// the current one is computed.  Loads go by 16-byte vectors of LV elements along
// whichever index is contiguous: one vector load when the step's block lies inside
// the operand and the operand allows aligned vectors, else LV element loads from
// indices clamped into the operand (the choice is uniform over the workgroup).
// Elements past m, n or K are zeroed when the step is written to LDS.  No load sits
// behind a per-element guard: a guarded load is a branch, and at each branch join
// the compiler waits for every load issued so far.
template <class TR, int BK>
__global__ __launch_bounds__(256) void gemm_small_kernel(GemmArgs g, const typename TR::T *__restrict__ A,
                                                         int64_t asr, int64_t ask, const typename TR::T *__restrict__ B,
                                                         int64_t bsr, int64_t bsk, int64_t K) {
  using T = typename TR::T;
  using acc_t = typename TR::acc_t;
  constexpr int VE = TR::VE, KF = TR::KF, SM = 32, KC = BK / KF;
  constexpr int LV = 16 / (int)sizeof(T);  // elements per 16-byte vector (== VE)
  constexpr int FV = KC * (SM / 16) * 64;  // fragment vectors per operand tile
  constexpr int NV = SM * BK / (LV * 256); // vectors per thread per operand
  static_assert(!TR::SPLIT && TR::PL == 1 && LV == VE, "plain MFMA types only");
  static_assert(BK % KF == 0 && NV >= 1 && (SM * BK) % (LV * 256) == 0, "k step must hold whole fragments");
  __shared__ v4u la[FV], lb[FV];
  T *las = reinterpret_cast<T *>(la);
  T *lbs = reinterpret_cast<T *>(lb);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6, wm = wave & 1, wn = wave >> 1;
  const int64_t m0 = (int64_t)blockIdx.x * SM, n0 = (int64_t)blockIdx.y * SM;
  const bool arows = asr == 1, brows = bsr == 1;
  // aligned 16-byte vectors along the contiguous index (its block starts are multiples of LV)
  const bool avec = ((uintptr_t)A % 16 == 0) && ((arows ? ask : asr) % LV == 0);
  const bool bvec = ((uintptr_t)B % 16 == 0) && ((brows ? bsk : bsr) % LV == 0);
  // first element (row r, k kk) of vector v of a step; the vector runs along r (rows
  // contiguous) or along kk
  auto vpos = [](bool rows, int v, int &r, int &kk) {
    r = rows ? (v % (SM / LV)) * LV : v / (BK / LV);
    kk = rows ? v / (SM / LV) : (v % (BK / LV)) * LV;
  };
  T va[NV][LV], vb[NV][LV];
  auto gather_one = [&](const T *X, int64_t sr, int64_t sk, bool rows, bool full, int64_t R, int64_t r0, int64_t k0,
                        T (&vx)[NV][LV]) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      int r, kk;
      vpos(rows, t + 256 * i, r, kk);
      if (full) {
        const v4u x = *reinterpret_cast<const v4u *>(X + (r0 + r) * sr + (k0 + kk) * sk);
        __builtin_memcpy(vx[i], &x, 16);
      } else {
#pragma unroll
        for (int e = 0; e < LV; ++e)
          vx[i][e] = X[imin64(r0 + r + (rows ? e : 0), R - 1) * sr + imin64(k0 + kk + (rows ? 0 : e), K - 1) * sk];
      }
    }
  };
  auto gather = [&](int64_t k0) {
    const bool kin = k0 + BK <= K;
    gather_one(A, asr, ask, arows, avec && kin && m0 + SM <= g.M, g.M, m0, k0, va);
    gather_one(B, bsr, bsk, brows, bvec && kin && n0 + SM <= g.N, g.N, n0, k0, vb);
  };
  auto frag_pos = [](int r, int kk) {  // element offset of (row r, k kk) in the fragment-ordered tile
    return (((kk / KF) * (SM / 16) + r / 16) * 64 + (r % 16) + 16 * ((kk % KF) / VE)) * VE + kk % VE;
  };
  auto store_one = [&](T *ls, bool rows, int64_t R, int64_t r0, int64_t k0, const T (&vx)[NV][LV]) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      int r, kk;
      vpos(rows, t + 256 * i, r, kk);
      if (rows) {  // LV rows of one k: LV fragment slots
#pragma unroll
        for (int e = 0; e < LV; ++e) ls[frag_pos(r + e, kk)] = (r0 + r + e < R && k0 + kk < K) ? vx[i][e] : T(0);
      } else {     // LV k values of one row: exactly one fragment vector
        T w[LV];
#pragma unroll
        for (int e = 0; e < LV; ++e) w[e] = (r0 + r < R && k0 + kk + e < K) ? vx[i][e] : T(0);
        v4u x;
        __builtin_memcpy(&x, w, 16);
        *reinterpret_cast<v4u *>(ls + frag_pos(r, kk)) = x;
      }
    }
  };
  acc_t acc = acc_t{};
  gather(0);
  for (int64_t k0 = 0; k0 < K; k0 += BK) {
    __syncthreads();  // the previous step's fragments are consumed
    store_one(las, arows, g.M, m0, k0, va);
    store_one(lbs, brows, g.N, n0, k0, vb);
    __syncthreads();
    if (k0 + BK < K) gather(k0 + BK);
#pragma unroll
    for (int kc = 0; kc < KC; ++kc) {
      const v4u fa = la[(kc * (SM / 16) + wm) * 64 + lane];
      const v4u fb = lb[(kc * (SM / 16) + wn) * 64 + lane];
      TR::mma(acc, &fa, &fb);
    }
  }
  const acc_t accs[1][1] = {{acc}};
  epilogue<TR, EPI_REAL, 1, 1>(accs, g, m0 + wm * 16, n0 + wn * 16, lane, 0);
}

// C := beta*C for the quick return (alpha == 0 or k == 0), never reading C when beta == 0.
template <typename T, bool CPLX>
__global__ void scale_c_kernel(T *C, int64_t m, int64_t n, int64_t ldc, double br, double bi) {
  int64_t total = m * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % m, j = idx / m;
    if constexpr (CPLX) {
      T *p = C + 2 * (i + j * ldc);
      if (br == 0.0 && bi == 0.0) { p[0] = T(0); p[1] = T(0); }
      else { T re = p[0], im = p[1]; p[0] = re * (T)br - im * (T)bi; p[1] = re * (T)bi + im * (T)br; }
    } else {
      T *p = C + i + j * ldc;
      *p = (br == 0.0) ? T(0) : (T)br * (*p);
    }
  }
}

// 3M combine: Re C = T1 - T2, Im C = T3 - T1 - T2, then C := alpha*(Re + i Im) + beta*C.
__global__ void combine_3m_kernel(const float *Tz, int64_t m, int64_t n, float *C, int64_t ldc,
                                  float ar, float ai, float br, float bi, int beta_zero) {
  int64_t total = m * n;
  int64_t zs = m * n;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    int64_t i = idx % m, j = idx / m;
    float t1 = Tz[idx], t2 = Tz[zs + idx], t3 = Tz[2 * zs + idx];
    float re = t1 - t2, im = t3 - t1 - t2;
    float *p = C + 2 * (i + j * ldc);
    float ore = ar * re - ai * im, oim = ar * im + ai * re;
    if (!beta_zero) {
      float cr = p[0], ci = p[1];
      ore += br * cr - bi * ci;
      oim += br * ci + bi * cr;
    }
    p[0] = ore;
    p[1] = oim;
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct TileShape { int bm, bn, bk; };

// Tile menus per element kind.  Index = tile id stored by the autotuner.
// WVM x WVN > 0 selects the direct-to-LDS kernel with that wave arrangement; 0 x 0
// the register-staged kernel (2 x 2 waves).
template <int BM_, int BN_, int BK_, int WVM_ = 0, int WVN_ = 0, int SCH_ = 0> struct Shape {
  static constexpr int bm = BM_, bn = BN_, bk = BK_, wvm = WVM_, wvn = WVN_, sch = SCH_;
};
template <class... S> struct ShapeList {};

// Tile menus per element kind; the index is the tile id the autotuner stores.  No
// entry spills registers: a 256 x 256 tile needs eight waves (4 waves would hold
// 64 accumulator tiles each), so the 4-wave entries are at most 128 x 256 for the
// split forms.
using MenuF64 = ShapeList<Shape<128, 128, 16>, Shape<64, 64, 16>, Shape<128, 128, 32>, Shape<128, 128, 16, 2, 2>,
                          Shape<128, 128, 16, 2, 4>, Shape<64, 64, 32, 2, 2>, Shape<128, 64, 32, 2, 2>,
                          Shape<256, 128, 16, 2, 4>, Shape<128, 128, 16, 2, 2, 2>, Shape<128, 128, 32, 2, 4, 2>>;
using MenuF32 = ShapeList<Shape<128, 128, 32>, Shape<64, 64, 32>, Shape<256, 128, 32>, Shape<256, 256, 32>,
                          Shape<128, 128, 64>, Shape<128, 256, 32, 2, 4>, Shape<256, 256, 32, 2, 4>,
                          Shape<128, 128, 64, 2, 2>, Shape<256, 128, 32, 2, 2>, Shape<256, 256, 32, 2, 4, 1>,
                          Shape<256, 256, 32, 4, 2, 1>, Shape<256, 256, 32, 2, 4, 2>, Shape<256, 256, 32, 4, 2, 2>>;
using MenuBF3 = ShapeList<Shape<128, 128, 32>, Shape<64, 64, 32>, Shape<256, 128, 32>, Shape<128, 256, 32>,
                          Shape<128, 128, 64>, Shape<128, 256, 32, 2, 4>, Shape<256, 256, 32, 2, 4>,
                          Shape<128, 128, 64, 2, 2>, Shape<256, 128, 32, 2, 2>, Shape<256, 256, 32, 2, 4, 1>,
                          Shape<256, 256, 32, 4, 2, 1>, Shape<256, 256, 32, 2, 4, 2>, Shape<256, 256, 32, 4, 2, 2>>;

template <class L> struct MenuTable;
template <class... S> struct MenuTable<ShapeList<S...>> {
  static const TileShape *t() {
    static const TileShape tab[] = {{S::bm, S::bn, S::bk}...};
    return tab;
  }
  static constexpr int n = sizeof...(S);
};

template <class TR> struct Menu;
template <> struct Menu<TF64> : MenuTable<MenuF64> { using list = MenuF64; };
template <> struct Menu<TF32> : MenuTable<MenuF32> { using list = MenuF32; };
template <> struct Menu<TBF3> : MenuTable<MenuBF3> { using list = MenuBF3; };
template <> struct Menu<TF16X3> : MenuTable<MenuBF3> { using list = MenuBF3; };

template <class TR, int BM, int BN, int BK, int WVM, int WVN, int SCH, int EPI>
static int launch_gemm(const GemmArgs &g, int zdim, hipStream_t s) {
  dim3 grid((unsigned)(g.tiles_m * g.tiles_n), 1, (unsigned)zdim);
  if constexpr (WVM > 0) {
    const int slot = timing_begin(TK_GEMM, s);
    hipLaunchKernelGGL((gemm_dma_kernel<TR, BM, BN, BK, WVM, WVN, SCH, EPI>), grid, dim3(64 * WVM * WVN), 0, s, g);
    CORAL_LAUNCH_CHECK();
    timing_end(slot, s);
  } else {
    constexpr int KC = BK / TR::KF;
    constexpr size_t lds_bytes = 2 * (size_t)KC * ((BM + BN) / 16) * TR::PL * 64 * 16;
    auto kfn = gemm_kernel<TR, BM, BN, BK, EPI>;
    static bool attr_set = false;
    if (!attr_set) {
      CORAL_HIP(hipFuncSetAttribute((const void *)kfn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
      attr_set = true;
    }
    const int slot = timing_begin(TK_GEMM, s);
    hipLaunchKernelGGL(kfn, grid, dim3(256), lds_bytes, s, g);
    CORAL_LAUNCH_CHECK();
    timing_end(slot, s);
  }
  return CORAL_OK;
}

// instantiate the kernel for every menu entry and launch entry `idx`
template <class TR, int EPI>
static int dispatch_tile(int, const GemmArgs &, int, hipStream_t, ShapeList<>) {
  set_last_error("gemm: tile id out of range");
  return CORAL_ERR_UNSUPPORTED;
}
template <class TR, int EPI, class S, class... R>
static int dispatch_tile(int idx, const GemmArgs &g, int zdim, hipStream_t s, ShapeList<S, R...>) {
  if (idx == 0) return launch_gemm<TR, S::bm, S::bn, S::bk, S::wvm, S::wvn, S::sch, EPI>(g, zdim, s);
  return dispatch_tile<TR, EPI>(idx - 1, g, zdim, s, ShapeList<R...>{});
}

template <class TR, int EPI>
static int launch_tile(int tile, const GemmArgs &g, int zdim, hipStream_t s) {
  if (tile < 0 || tile >= Menu<TR>::n) tile = 0;
  return dispatch_tile<TR, EPI>(tile, g, zdim, s, typename Menu<TR>::list{});
}

template <class TR, int BR, int BK, class F>
static int launch_pack_t(v4u *dst, int64_t R, int64_t K, F f, hipStream_t s) {
  if constexpr (BK % TR::KF != 0) {
    set_last_error("pack: k tile %d is not a multiple of %d", BK, TR::KF);
    return CORAL_ERR_UNSUPPORTED;
  } else {
    int64_t RT = cdiv(R, BR), KT = cdiv(K, BK);
    constexpr int UPT = (BK / TR::KF) * (BR / 16) * 64;
    int64_t units = RT * KT * UPT;
    hipLaunchKernelGGL((pack_kernel<TR, BR, BK, F>), dim3((unsigned)cdiv(units, 256)), dim3(256), 0, s, dst, R, K,
                       KT, units, f);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
}

// Job for one real operand: pack (fragment-ordered tiles of br x bk) or absmax.
template <typename T>
static RealJob<T> real_job(const FetchReal<T> &f, int64_t R, int64_t K, int br, int bk, v4u *dst, uint32_t *mx,
                           bool absmax) {
  constexpr int LV = 16 / sizeof(T);
  RealJob<T> j{};
  j.dst = dst;
  j.p = f.p;
  j.sr = f.sr;
  j.sk = f.sk;
  j.R = R;
  j.K = K;
  j.KT = cdiv(K, bk);
  j.Rp = cdiv(R, br) * br;
  j.Kp = j.KT * bk;
  j.mx = mx;
  j.br = br;
  j.rows_contig = f.sr == 1;
  const int64_t stride = j.rows_contig ? f.sk : f.sr;
  j.vec = ((uintptr_t)f.p % 16 == 0) && (stride % LV == 0);
  if (absmax) {  // shapes of absmax_real_kernel
    j.rblocks = j.rows_contig ? cdiv(R, 256) : cdiv(R, 4);
    j.kblocks = j.rows_contig ? cdiv(K, 256) : cdiv(K, 8192);
  } else {       // shapes of pack_real_kernel
    j.rblocks = j.rows_contig ? cdiv(j.Rp, 64 * LV) : cdiv(j.Rp, 64);
    j.kblocks = j.rows_contig ? j.KT : cdiv(j.Kp, 32 * LV);
  }
  j.blocks = j.rblocks * j.kblocks;
  return j;
}

template <class TR, int BK>
static int launch_pack_real_bk(const RealJob<typename TR::T> &ja, const RealJob<typename TR::T> &jb, bool scan,
                               const uint32_t *gate, hipStream_t s) {
  if constexpr (BK % TR::KF != 0) {
    set_last_error("pack: k tile %d is not a multiple of %d", BK, TR::KF);
    return CORAL_ERR_UNSUPPORTED;
  } else {
    const dim3 grid((unsigned)(ja.blocks + jb.blocks));
    if constexpr (TR::SPLIT) {
      if (scan) {
        hipLaunchKernelGGL((pack_real_kernel<TR, BK, true>), grid, dim3(256), 0, s, ja, jb, gate);
        CORAL_LAUNCH_CHECK();
        return CORAL_OK;
      }
    }
    hipLaunchKernelGGL((pack_real_kernel<TR, BK, false>), grid, dim3(256), 0, s, ja, jb, gate);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
}

// Packs op(A) and op(B) in one launch.
template <class TR>
static int launch_pack_real(const TileShape &ts, v4u *pa, v4u *pb, int64_t M, int64_t N, int64_t K,
                            const FetchReal<typename TR::T> &fa, const FetchReal<typename TR::T> &fb,
                            uint32_t *rexp, uint32_t *cexp, bool scan, const uint32_t *gate, hipStream_t s) {
  const auto ja = real_job(fa, M, K, ts.bm, ts.bk, pa, rexp, false);
  const auto jb = real_job(fb, N, K, ts.bn, ts.bk, pb, cexp, false);
  if (ts.bk == 16) return launch_pack_real_bk<TR, 16>(ja, jb, scan, gate, s);
  if (ts.bk == 32) return launch_pack_real_bk<TR, 32>(ja, jb, scan, gate, s);
  if (ts.bk == 64) return launch_pack_real_bk<TR, 64>(ja, jb, scan, gate, s);
  set_last_error("pack: no kernel for k tile %d", ts.bk);
  return CORAL_ERR_UNSUPPORTED;
}

// Job for one complex 3M operand (pack_cplx3m_kernel); dst holds three planes of
// dst_zstride vectors, mx three rows of maxima mx_zstride apart.
static Cplx3MJob cplx3m_job(const FetchCplx3M<float> &f, int64_t R, int64_t K, int br, int bk, v4u *dst,
                            int64_t dst_zstride, uint32_t *mx, int64_t mx_zstride) {
  Cplx3MJob j{};
  j.dst = dst;
  j.dst_zstride = dst_zstride;
  j.p = f.p;
  j.sr = f.sr;
  j.sk = f.sk;
  j.R = R;
  j.K = K;
  j.KT = cdiv(K, bk);
  j.Rp = cdiv(R, br) * br;
  j.Kp = j.KT * bk;
  j.mx = mx;
  j.mx_zstride = mx_zstride;
  j.br = br;
  j.rows_contig = f.sr == 1;
  j.conj = f.conj;
  const int64_t stride = j.rows_contig ? f.sk : f.sr;  // in complex elements
  j.vec = ((uintptr_t)f.p % 16 == 0) && (stride % 2 == 0);
  j.rblocks = j.rows_contig ? cdiv(j.Rp, 128) : cdiv(j.Rp, 64);
  j.kblocks = j.rows_contig ? j.KT : cdiv(j.Kp, 64);
  j.blocks = j.rblocks * j.kblocks;
  return j;
}

// mode: 0 = pack scaled by mx (if set), 1 = pack unscaled and take the maxima,
// 2 = take the maxima only.
template <class TR, int BK>
static int launch_pack_3m_bk(const Cplx3MJob &ja, const Cplx3MJob &jb, int mode, const uint32_t *gate,
                             hipStream_t s) {
  if constexpr (BK % TR::KF != 0) {
    set_last_error("pack: k tile %d is not a multiple of %d", BK, TR::KF);
    return CORAL_ERR_UNSUPPORTED;
  } else {
    const dim3 grid((unsigned)(ja.blocks + jb.blocks));
    if (mode == 1)
      hipLaunchKernelGGL((pack_cplx3m_kernel<TR, BK, true, true>), grid, dim3(256), 0, s, ja, jb, gate);
    else if (mode == 2)
      hipLaunchKernelGGL((pack_cplx3m_kernel<TR, BK, true, false>), grid, dim3(256), 0, s, ja, jb, gate);
    else
      hipLaunchKernelGGL((pack_cplx3m_kernel<TR, BK, false, true>), grid, dim3(256), 0, s, ja, jb, gate);
    CORAL_LAUNCH_CHECK();
    return CORAL_OK;
  }
}

// Packs the three planes of op(A) and op(B) in one launch (or takes their maxima).
template <class TR>
static int launch_pack_3m(const TileShape &ts, const Cplx3MJob &ja, const Cplx3MJob &jb, int mode,
                          const uint32_t *gate, hipStream_t s) {
  if (ts.bk == 32) return launch_pack_3m_bk<TR, 32>(ja, jb, mode, gate, s);
  if (ts.bk == 64) return launch_pack_3m_bk<TR, 64>(ja, jb, mode, gate, s);
  set_last_error("pack: no 3M kernel for k tile %d", ts.bk);
  return CORAL_ERR_UNSUPPORTED;
}

// Row maxima of op(A) (m rows) and op(B) (n rows) into ex[0, m) and ex[m, m+n).
static int launch_absmax_real(const FetchReal<float> &fa, const FetchReal<float> &fb, int64_t m, int64_t n,
                              int64_t k, uint32_t *ex, hipStream_t s) {
  const auto ja = real_job(fa, m, k, 16, 16, nullptr, ex, true);
  const auto jb = real_job(fb, n, k, 16, 16, nullptr, ex + m, true);
  hipLaunchKernelGGL(absmax_real_kernel, dim3((unsigned)(ja.blocks + jb.blocks)), dim3(256), 0, s, ja, jb);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

template <class TR, class F>
static int launch_pack(const TileShape &ts, bool is_a, v4u *dst, int64_t R, int64_t K, F f, hipStream_t s) {
  int br = is_a ? ts.bm : ts.bn;
  if (ts.bk == 16) {
    if (br == 64) return launch_pack_t<TR, 64, 16>(dst, R, K, f, s);
    if (br == 128) return launch_pack_t<TR, 128, 16>(dst, R, K, f, s);
    return launch_pack_t<TR, 256, 16>(dst, R, K, f, s);
  } else if (ts.bk == 32) {
    if (br == 64) return launch_pack_t<TR, 64, 32>(dst, R, K, f, s);
    if (br == 128) return launch_pack_t<TR, 128, 32>(dst, R, K, f, s);
    return launch_pack_t<TR, 256, 32>(dst, R, K, f, s);
  } else {
    if (br == 64) return launch_pack_t<TR, 64, 64>(dst, R, K, f, s);
    if (br == 128) return launch_pack_t<TR, 128, 64>(dst, R, K, f, s);
    return launch_pack_t<TR, 256, 64>(dst, R, K, f, s);
  }
}

static int zero_u32(uint32_t *p, int64_t count, hipStream_t s) {
  hipLaunchKernelGGL(zero_u32_kernel, dim3((unsigned)std::min<int64_t>(cdiv(count, 256), 1024)), dim3(256), 0, s, p,
                     count);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

template <class TR>
static size_t packed_bytes(int64_t R, int64_t K, int br, int bk) {
  int64_t Rp = cdiv(R, br) * br, Kp = cdiv(K, bk) * bk;
  return (size_t)Rp * (size_t)Kp * (16 / TR::VE) * TR::PL;
}

// Default tile choice when the autotuner has no entry: large tiles only when the
// grid still fills the chip.
template <class TR>
static int default_tile(int64_t M, int64_t N, int64_t K) {
  (void)K;
  const TileShape *t = Menu<TR>::t();
  const int cus = num_cus();
  auto tiles = [&](const TileShape &s) { return cdiv(M, s.bm) * cdiv(N, s.bn); };
  if constexpr (std::is_same<TR, TF64>::value) {
    // 128x128x16, 8 waves, direct-to-LDS: 70.0 TFLOP/s at 8192^3 (0.89 of the FP64
    // MFMA peak) vs 61.1 for the register-staged 64x64 tile (id 1), which still
    // serves grids with fewer 128x128 tiles than CUs
    if (tiles(t[4]) >= cus) return 4;
    return 1;
  } else {
    // 256x256 tiles, 8 waves, direct-to-LDS loads: fewest LDS bytes per MFMA.  At
    // 8192^3 (warm clocks): BF16x3 492 TFLOP/s with the row-pipelined schedule
    // (id 11) vs 448 compiler-scheduled (id 6) and 361 register-staged (id 3);
    // plain FP32 141 vs 134 (id 0).
    constexpr int big = TR::SPLIT ? 11 : 6;
    if (tiles(t[big]) >= cus) return big;
  }
  if (tiles(t[0]) >= cus) return 0;
  return 1;
}


// Shared driver for the real and complex-4M products.
//   kind: 0 = real, 1 = complex 4M
// gate (split real products only, may be null): rexp, cexp and *gate hold zeros; the
// operands are packed unscaled while their row maxima are taken, split_gate_kernel
// decides, and the scaled re-pack and the epilogue's rescale run only if it said so.
template <class TR, class FA, class FB>
static int run_packed(const TileShape &ts, int epi, int64_t M, int64_t N, int64_t K, FA fa, FB fb,
                      const uint32_t *rexp, const uint32_t *cexp, void *C, int64_t ldc, double ar, double ai,
                      double br, double bi, int tile, hipStream_t s, uint32_t *gate = nullptr) {
  const int64_t KT = cdiv(K, ts.bk);
  void *pa = nullptr, *pb = nullptr;
  CORAL_TRY(workspace(s, WS_PACK_A, packed_bytes<TR>(M, K, ts.bm, ts.bk), &pa));
  CORAL_TRY(workspace(s, WS_PACK_B, packed_bytes<TR>(N, K, ts.bn, ts.bk), &pb));
  if constexpr (std::is_same<FA, FetchReal<typename TR::T>>::value) {
    if (gate) {
      CORAL_TRY(launch_pack_real<TR>(ts, (v4u *)pa, (v4u *)pb, M, N, K, fa, fb, (uint32_t *)rexp, (uint32_t *)cexp,
                                     true, nullptr, s));
      hipLaunchKernelGGL(split_gate_kernel, dim3((unsigned)cdiv(M + N, 1024)), dim3(1024), 0, s, rexp, M, cexp, N,
                         gate);
      CORAL_LAUNCH_CHECK();
    }
    CORAL_TRY(launch_pack_real<TR>(ts, (v4u *)pa, (v4u *)pb, M, N, K, fa, fb, (uint32_t *)rexp, (uint32_t *)cexp,
                                   false, gate, s));
  } else {
    CORAL_TRY(launch_pack<TR>(ts, true, (v4u *)pa, M, K, fa, s));
    CORAL_TRY(launch_pack<TR>(ts, false, (v4u *)pb, N, K, fb, s));
  }
  GemmArgs g{};
  g.Ap = (const v4u *)pa;
  g.Bp = (const v4u *)pb;
  g.M = M;
  g.N = N;
  g.KT = KT;
  g.tiles_m = (int)cdiv(M, ts.bm);
  g.tiles_n = (int)cdiv(N, ts.bn);
  g.C = C;
  g.ldc = ldc;
  g.ar = ar; g.ai = ai; g.br = br; g.bi = bi;
  g.beta_zero = (br == 0.0 && bi == 0.0);
  g.row_exp = rexp;
  g.col_exp = cexp;
  g.gate = gate;
  if (epi == EPI_CPLX) return launch_tile<TR, EPI_CPLX>(tile, g, 1, s);
  return launch_tile<TR, EPI_REAL>(tile, g, 1, s);
}

template <typename T>
static void strides_a(int op, int64_t lda, int64_t &sr, int64_t &sk) {
  if (op == CORAL_NO_TRANSPOSE) { sr = 1; sk = lda; } else { sr = lda; sk = 1; }
}
template <typename T>
static void strides_b(int op, int64_t ldb, int64_t &sr, int64_t &sk) {
  // rows of the packed B operand are columns j of op(B)
  if (op == CORAL_NO_TRANSPOSE) { sr = ldb; sk = 1; } else { sr = 1; sk = ldb; }
}

// argument validation shared by all gemm entry points (CORAL argument order:
// op_a=1 op_b=2 m=3 n=4 k=5 alpha=6 a=7 lda=8 b=9 ldb=10 beta=11 c=12 ldc=13)
static int check_gemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const void *a,
                      int64_t lda, const void *b, int64_t ldb, const void *c, int64_t ldc, bool need_ab) {
  if (!valid_op(op_a)) return -1;
  if (!valid_op(op_b)) return -2;
  if (m < 0) return -3;
  if (n < 0) return -4;
  if (k < 0) return -5;
  int64_t arows = (op_a == CORAL_NO_TRANSPOSE) ? m : k;
  int64_t brows = (op_b == CORAL_NO_TRANSPOSE) ? k : n;
  if (need_ab && !a) return -7;
  if (lda < imax64(1, arows)) return -8;
  if (need_ab && !b) return -9;
  if (ldb < imax64(1, brows)) return -10;
  if (!c && m > 0 && n > 0) return -12;
  if (ldc < imax64(1, m)) return -13;
  return CORAL_OK;
}

template <typename T, bool CPLX>
static int quick_scale(T *c, int64_t m, int64_t n, int64_t ldc, double br, double bi, hipStream_t s) {
  if (!CPLX && br == 1.0) return CORAL_OK;
  if (CPLX && br == 1.0 && bi == 0.0) return CORAL_OK;
  int64_t blocks = std::min<int64_t>(cdiv(m * n, 256), 4096);
  hipLaunchKernelGGL((scale_c_kernel<T, CPLX>), dim3((unsigned)blocks), dim3(256), 0, s, c, m, n, ldc, br, bi);
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

// k step of gemm_small_kernel: 32 for K up to 32, else 64 (SGEMM 512^3 with
// alpha = 1, beta = 0.5: 12.4 / 11.8 / 12.4 us per call with 32 / 64 / 128).
// CORAL_SMALL_BK = 32 or 64 overrides, for measurement and tests.
static int small_bk(int64_t k) {
  static const int forced = [] {
    const char *e = getenv("CORAL_SMALL_BK");
    return e ? atoi(e) : 0;
  }();
  if (forced == 32 || forced == 64) return forced;
  return k <= 32 ? 32 : 64;
}

template <class TR, typename T>
static int real_gemm(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, T alpha,
                     const T *a, int64_t lda, const T *b, int64_t ldb, T beta, T *c, int64_t ldc,
                     coral_stream_t stream) {
  bool trivial = (alpha == T(0) || k == 0);
  CORAL_TRY(check_gemm(op_a, op_b, m, n, k, a, lda, b, ldb, c, ldc, !trivial && m > 0 && n > 0));
  if (m == 0 || n == 0) return CORAL_OK;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (trivial) return quick_scale<T, false>(c, m, n, ldc, (double)beta, 0.0, s);
  int tile = tuned_tile(routine, op_a, op_b, m, n, k);
  FetchReal<T> fa{a, 0, 0}, fb{b, 0, 0};
  strides_a<T>(op_a, lda, fa.sr, fa.sk);
  strides_b<T>(op_b, ldb, fb.sr, fb.sk);
  if constexpr (!TR::SPLIT) {
    // small problems: too few 64x64 tiles to fill the chip; one unpacked launch
    if (tile < 0 && cdiv(m, 64) * cdiv(n, 64) < num_cus()) {
      GemmArgs g{};
      g.M = m;
      g.N = n;
      g.C = c;
      g.ldc = ldc;
      g.ar = (double)alpha;
      g.br = (double)beta;
      g.beta_zero = beta == T(0);
      const int slot = timing_begin(TK_GEMM, s);
      const dim3 grid((unsigned)cdiv(m, 32), (unsigned)cdiv(n, 32));
      const int bk = small_bk(k);
      if (bk == 32)
        hipLaunchKernelGGL((gemm_small_kernel<TR, 32>), grid, dim3(256), 0, s, g, a, fa.sr, fa.sk, b, fb.sr, fb.sk, k);
      else
        hipLaunchKernelGGL((gemm_small_kernel<TR, 64>), grid, dim3(256), 0, s, g, a, fa.sr, fa.sk, b, fb.sr, fb.sk, k);
      CORAL_LAUNCH_CHECK();
      timing_end(slot, s);
      return CORAL_OK;
    }
  }
  if (tile < 0 || tile >= Menu<TR>::n) tile = default_tile<TR>(m, n, k);
  const TileShape ts = Menu<TR>::t()[tile];
  if constexpr (TR::SPLIT) {
    uint32_t *ex = nullptr;
    CORAL_TRY(workspace(s, WS_AUX0, sizeof(uint32_t) * (size_t)(m + n + 1), (void **)&ex));
    CORAL_TRY(zero_u32(ex, m + n + 1, s));  // row maxima, column maxima, gate
    if constexpr (TR::SCALE_EXP == 0) {
      // BF16 halves: FP32's exponent range, so the scaling pass is needed only for
      // extreme rows; the maxima come from the packing pass itself
      return run_packed<TR>(ts, EPI_REAL, m, n, k, fa, fb, ex, ex + m, c, ldc, (double)alpha, 0.0, (double)beta,
                            0.0, tile, s, ex + m + n);
    } else {
      CORAL_TRY(launch_absmax_real(fa, fb, m, n, k, ex, s));
      return run_packed<TR>(ts, EPI_REAL, m, n, k, fa, fb, ex, ex + m, c, ldc, (double)alpha, 0.0, (double)beta,
                            0.0, tile, s);
    }
  } else {
    return run_packed<TR>(ts, EPI_REAL, m, n, k, fa, fb, nullptr, nullptr, c, ldc, (double)alpha, 0.0,
                          (double)beta, 0.0, tile, s);
  }
}

template <class TR, typename T>
static int complex_gemm_4m(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                           const T *alpha, const T *a, int64_t lda, const T *b, int64_t ldb, const T *beta,
                           T *c, int64_t ldc, coral_stream_t stream) {
  if (!alpha) return -6;
  if (!beta) return -11;
  bool trivial = (alpha[0] == T(0) && alpha[1] == T(0)) || k == 0;
  CORAL_TRY(check_gemm(op_a, op_b, m, n, k, a, lda, b, ldb, c, ldc, !trivial && m > 0 && n > 0));
  if (m == 0 || n == 0) return CORAL_OK;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (trivial) return quick_scale<T, true>(c, m, n, ldc, (double)beta[0], (double)beta[1], s);
  // real product: (m) x (2n) x (2k)
  int tile = tuned_tile(routine, op_a, op_b, m, n, k);
  if (tile < 0 || tile >= Menu<TR>::n) tile = default_tile<TR>(m, 2 * n, 2 * k);
  const TileShape ts = Menu<TR>::t()[tile];
  FetchCplxA4M<T> fa{a, 0, 0, k, op_a == CORAL_CONJUGATE_TRANSPOSE};
  FetchCplxB4M<T> fb{b, 0, 0, k, op_b == CORAL_CONJUGATE_TRANSPOSE};
  strides_a<T>(op_a, lda, fa.sr, fa.sk);
  strides_b<T>(op_b, ldb, fb.sr, fb.sk);
  return run_packed<TR>(ts, EPI_CPLX, m, 2 * n, 2 * k, fa, fb, nullptr, nullptr, c, ldc, (double)alpha[0],
                        (double)alpha[1], (double)beta[0], (double)beta[1], tile, s);
}

template <class TR>
static int cgemm_3m(const char *routine, int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k,
                    const float *alpha, const float *a, int64_t lda, const float *b, int64_t ldb, const float *beta,
                    float *c, int64_t ldc, coral_stream_t stream) {
  if (!alpha) return -6;
  if (!beta) return -11;
  bool trivial = (alpha[0] == 0.f && alpha[1] == 0.f) || k == 0;
  CORAL_TRY(check_gemm(op_a, op_b, m, n, k, a, lda, b, ldb, c, ldc, !trivial && m > 0 && n > 0));
  if (m == 0 || n == 0) return CORAL_OK;
  CORAL_TRY(ensure_init());
  hipStream_t s = resolve(stream);
  if (trivial) return quick_scale<float, true>(c, m, n, ldc, beta[0], beta[1], s);
  int tile = tuned_tile(routine, op_a, op_b, m, n, k);
  if (tile < 0 || tile >= Menu<TR>::n) tile = default_tile<TR>(m, n, k);
  const TileShape ts = Menu<TR>::t()[tile];
  FetchCplx3M<float> fa{a, 0, 0, op_a == CORAL_CONJUGATE_TRANSPOSE};
  FetchCplx3M<float> fb{b, 0, 0, op_b == CORAL_CONJUGATE_TRANSPOSE};
  strides_a<float>(op_a, lda, fa.sr, fa.sk);
  strides_b<float>(op_b, ldb, fb.sr, fb.sk);
  // row maxima: 3 x m for A, 3 x n for B, then the gate word; products T_z (3 x m x n fp32)
  uint32_t *ex = nullptr;
  CORAL_TRY(workspace(s, WS_AUX0, sizeof(uint32_t) * (3 * (size_t)(m + n) + 1), (void **)&ex));
  uint32_t *exa = ex, *exb = ex + 3 * m, *gate = nullptr;
  CORAL_TRY(zero_u32(ex, 3 * (m + n) + 1, s));
  size_t abytes = packed_bytes<TR>(m, k, ts.bm, ts.bk), bbytes = packed_bytes<TR>(n, k, ts.bn, ts.bk);
  void *pa = nullptr, *pb = nullptr, *tz = nullptr;
  CORAL_TRY(workspace(s, WS_PACK_A, 3 * abytes, &pa));
  CORAL_TRY(workspace(s, WS_PACK_B, 3 * bbytes, &pb));
  CORAL_TRY(workspace(s, WS_AUX1, 3 * sizeof(float) * (size_t)m * (size_t)n, &tz));
  const Cplx3MJob ja = cplx3m_job(fa, m, k, ts.bm, ts.bk, (v4u *)pa, (int64_t)(abytes / 16), exa, m);
  const Cplx3MJob jb = cplx3m_job(fb, n, k, ts.bn, ts.bk, (v4u *)pb, (int64_t)(bbytes / 16), exb, n);
  if constexpr (TR::SCALE_EXP == 0) {
    // BF16 halves: unscaled planes and their maxima in one pass, scaled re-pack only
    // if the gate is set (as for the real split form)
    gate = ex + 3 * (m + n);
    CORAL_TRY(launch_pack_3m<TR>(ts, ja, jb, 1, nullptr, s));
    hipLaunchKernelGGL(split_gate_kernel, dim3((unsigned)cdiv(3 * (m + n), 1024)), dim3(1024), 0, s, exa, 3 * m, exb,
                       3 * n, gate);
    CORAL_LAUNCH_CHECK();
    CORAL_TRY(launch_pack_3m<TR>(ts, ja, jb, 0, gate, s));
  } else {  // FP16 halves need the scaling: maxima first, then the scaled planes
    CORAL_TRY(launch_pack_3m<TR>(ts, ja, jb, 2, nullptr, s));
    CORAL_TRY(launch_pack_3m<TR>(ts, ja, jb, 0, nullptr, s));
  }
  GemmArgs g{};
  g.Ap = (const v4u *)pa;
  g.Bp = (const v4u *)pb;
  g.a_zstride = (int64_t)(abytes / 16);
  g.b_zstride = (int64_t)(bbytes / 16);
  g.M = m;
  g.N = n;
  g.KT = cdiv(k, ts.bk);
  g.tiles_m = (int)cdiv(m, ts.bm);
  g.tiles_n = (int)cdiv(n, ts.bn);
  g.C = tz;
  g.ldc = m;
  g.c_zstride = m * n;
  g.ar = 1.0;
  g.beta_zero = 1;
  g.row_exp = exa;
  g.col_exp = exb;
  g.re_zstride = m;
  g.ce_zstride = n;
  g.gate = gate;
  CORAL_TRY((launch_tile<TR, EPI_STORE_Z>(tile, g, 3, s)));
  int64_t blocks = std::min<int64_t>(cdiv(m * n, 256), 8192);
  hipLaunchKernelGGL(combine_3m_kernel, dim3((unsigned)blocks), dim3(256), 0, s, (const float *)tz, m, n, c, ldc,
                     alpha[0], alpha[1], beta[0], beta[1], (int)(beta[0] == 0.f && beta[1] == 0.f));
  CORAL_LAUNCH_CHECK();
  return CORAL_OK;
}

}  // namespace gemm

int gemm_menu_size(const char *routine) {
  if (!strcmp(routine, "dgemm") || !strcmp(routine, "zgemm")) return gemm::Menu<gemm::TF64>::n;
  if (!strcmp(routine, "sgemm") || !strcmp(routine, "cgemm")) return gemm::Menu<gemm::TF32>::n;
  if (!strcmp(routine, "sgemm_bf16x3") || !strcmp(routine, "cgemm_3m_bf16x3")) return gemm::Menu<gemm::TBF3>::n;
  if (!strcmp(routine, "sgemm_fp16x3") || !strcmp(routine, "cgemm_3m_fp16x3")) return gemm::Menu<gemm::TF16X3>::n;
  return 0;
}

}  // namespace coral

using namespace coral;
using namespace coral::gemm;

extern "C" {

int coral_sgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha, const float *a,
                int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc, coral_stream_t stream) {
  return real_gemm<TF32, float>("sgemm", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_dgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, double alpha, const double *a,
                int64_t lda, const double *b, int64_t ldb, double beta, double *c, int64_t ldc,
                coral_stream_t stream) {
  return real_gemm<TF64, double>("dgemm", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_sgemm_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha, const float *a,
                       int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc,
                       coral_stream_t stream) {
  return real_gemm<TBF3, float>("sgemm_bf16x3", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_cgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2], const float *a,
                int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c, int64_t ldc,
                coral_stream_t stream) {
  return complex_gemm_4m<TF32, float>("cgemm", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_zgemm(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const double alpha[2],
                const double *a, int64_t lda, const double *b, int64_t ldb, const double beta[2], double *c,
                int64_t ldc, coral_stream_t stream) {
  return complex_gemm_4m<TF64, double>("zgemm", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_cgemm_3m_bf16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2],
                          const float *a, int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c,
                          int64_t ldc, coral_stream_t stream) {
  return cgemm_3m<TBF3>("cgemm_3m_bf16x3", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_sgemm_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, float alpha, const float *a,
                       int64_t lda, const float *b, int64_t ldb, float beta, float *c, int64_t ldc,
                       coral_stream_t stream) {
  return real_gemm<TF16X3, float>("sgemm_fp16x3", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}
int coral_cgemm_3m_fp16x3(int32_t op_a, int32_t op_b, int64_t m, int64_t n, int64_t k, const float alpha[2],
                          const float *a, int64_t lda, const float *b, int64_t ldb, const float beta[2], float *c,
                          int64_t ldc, coral_stream_t stream) {
  return cgemm_3m<TF16X3>("cgemm_3m_fp16x3", op_a, op_b, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}

}  // extern "C"
