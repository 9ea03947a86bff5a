// MFMA element traits shared by the GEMM kernels (gemm.hip, batched.hip).
//
// A "fragment vector" is 16 bytes per lane.  Lane l of a 16x16 MFMA operand holds
// row (l % 16) and k-values (l / 16) * VE + t, t < VE, with VE = 16 / sizeof(elem);
// A and B use the same (lane, t) -> k map, so every 16-byte vector feeds VE 16x16x4
// MFMAs (f32/f64) or one 16x16x32 BF16/FP16 MFMA (hi and lo planes: PL = 2).
// The 16x16 accumulator holds column l % 16 in lane l; register r holds row
// 4*(l/16)+r for f32/bf16 MFMA and row l/16+4*r for f64 MFMA (ROWS_INTERLEAVED).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace coral {
namespace gemm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef uint32_t v4u __attribute__((ext_vector_type(4)));  // one 16-byte fragment vector

// ---------------------------------------------------------------------------
// element traits
// ---------------------------------------------------------------------------
struct TF64 {  // FP64 MFMA 16x16x4, DGEMM / ZGEMM
  using T = double;
  using acc_t = f64x4;
  static constexpr int VE = 2, PL = 1, KF = 8, NMMA = 2;  // NMMA: MFMA instructions per mma()
  static constexpr bool SPLIT = false;
  static constexpr bool ROWS_INTERLEAVED = true;
  __device__ static __forceinline__ void mma(acc_t &c, const v4u *a, const v4u *b) {
    f64x2 av = __builtin_bit_cast(f64x2, a[0]);
    f64x2 bv = __builtin_bit_cast(f64x2, b[0]);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(av[0], bv[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(av[1], bv[1], c, 0, 0, 0);
  }
};

struct TF32 {  // FP32 MFMA 16x16x4, plain SGEMM / CGEMM 4M
  using T = float;
  using acc_t = f32x4;
  static constexpr int VE = 4, PL = 1, KF = 16, NMMA = 4;
  static constexpr bool SPLIT = false;
  static constexpr bool ROWS_INTERLEAVED = false;
  __device__ static __forceinline__ void mma(acc_t &c, const v4u *a, const v4u *b) {
    f32x4 av = __builtin_bit_cast(f32x4, a[0]);
    f32x4 bv = __builtin_bit_cast(f32x4, b[0]);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[0], bv[0], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[1], bv[1], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[2], bv[2], c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(av[3], bv[3], c, 0, 0, 0);
  }
};

// Split forms: x = hi + lo + r with hi = rn(x), lo = rn(x - hi) in a 16-bit type;
// x - hi is exact in fp32.  The operand is first scaled by a power of two per row
// of op(A) / column of op(B) so that its largest magnitude has exponent SCALE_EXP.
struct TBF3 {  // BF16 hi/lo planes (8+8 significand bits), three 16x16x32 BF16 MFMAs, FP32 accumulation
  using T = float;
  using acc_t = f32x4;
  static constexpr int VE = 8, PL = 2, KF = 32, NMMA = 3;
  static constexpr bool SPLIT = true;
  static constexpr bool ROWS_INTERLEAVED = false;
  static constexpr int SCALE_EXP = 0;  // BF16 has FP32's exponent range: scale only to centre it
  __device__ static __forceinline__ void split(float x, uint32_t &hi, uint32_t &lo) {
    const __bf16 h = (__bf16)x;  // v_cvt_pk_bf16_f32: round to nearest even
    const __bf16 l = (__bf16)(x - (float)h);
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, l);
  }
  __device__ static __forceinline__ void mma(acc_t &c, const v4u *a, const v4u *b) {
    bf16x8 ah = __builtin_bit_cast(bf16x8, a[0]);
    bf16x8 al = __builtin_bit_cast(bf16x8, a[1]);
    bf16x8 bh = __builtin_bit_cast(bf16x8, b[0]);
    bf16x8 bl = __builtin_bit_cast(bf16x8, b[1]);
    // small cross terms first, then the dominant hi*hi product
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, c, 0, 0, 0);
  }
};

// FP16 hi/lo planes (11+11 significand bits), three 16x16x32 FP16 MFMAs, FP32
// accumulation.  FP16's range is [2^-24, 65504]: every row/column is scaled so its
// largest magnitude lies in [2^14, 2^15), which keeps hi finite after rounding
// (rn(x) <= 2^15) and leaves lo normal down to |x| >= 2^-18 * max; below that lo
// is subnormal with absolute spacing 2^-24, i.e. 2^-39 of the row maximum.
struct TF16X3 {
  using T = float;
  using acc_t = f32x4;
  static constexpr int VE = 8, PL = 2, KF = 32, NMMA = 3;
  static constexpr bool SPLIT = true;
  static constexpr bool ROWS_INTERLEAVED = false;
  static constexpr int SCALE_EXP = 14;
  __device__ static __forceinline__ void split(float x, uint32_t &hi, uint32_t &lo) {
    const _Float16 h = (_Float16)x;  // v_cvt_f16_f32: round to nearest even
    const _Float16 l = (_Float16)(x - (float)h);
    hi = __builtin_bit_cast(uint16_t, h);
    lo = __builtin_bit_cast(uint16_t, l);
  }
  __device__ static __forceinline__ void mma(acc_t &c, const v4u *a, const v4u *b) {
    f16x8 ah = __builtin_bit_cast(f16x8, a[0]);
    f16x8 al = __builtin_bit_cast(f16x8, a[1]);
    f16x8 bh = __builtin_bit_cast(f16x8, b[0]);
    f16x8 bl = __builtin_bit_cast(f16x8, b[1]);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, c, 0, 0, 0);
  }
};

}  // namespace gemm
}  // namespace coral
