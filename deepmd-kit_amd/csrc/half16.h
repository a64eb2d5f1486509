// 16-bit MFMA operand formats of the mixed-precision kernels (FP32
// accumulation in both):
//   Bf16 -- bfloat16, 8-bit significand, FP32's exponent range;
//   F16  -- IEEE half, 11-bit significand (8x finer rounding), range
//           6.1e-5 .. 65504 (normal).
// Values cross memory and registers as raw 16-bit words packed in pairs
// (unsigned), so kernels are written once against this interface.
#pragma once

#include "dpk_common.h"

namespace dpk {
namespace h16 {

typedef __attribute__((ext_vector_type(4))) float v4f;
typedef __attribute__((ext_vector_type(2))) float v2f;
typedef __attribute__((ext_vector_type(4))) unsigned v4u;
typedef __attribute__((ext_vector_type(2))) unsigned v2u;
typedef __attribute__((ext_vector_type(4))) short v4s;

struct Bf16 {
  static constexpr int kPrec = DPK_PREC_BF16;
  static constexpr bool kSplit = false;
  typedef __attribute__((ext_vector_type(2))) __bf16 v2;
  typedef __attribute__((ext_vector_type(8))) __bf16 v8;
  // two floats -> one packed pair, round to nearest even (v_cvt_pk_bf16_f32)
  __device__ static __forceinline__ unsigned pack2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a, b}, v2));
  }
  __device__ static __forceinline__ float lo(unsigned u) { return __uint_as_float(u << 16); }
  __device__ static __forceinline__ float hi(unsigned u) { return __uint_as_float(u & 0xffff0000u); }
  __device__ static __forceinline__ unsigned short from_f(float v) {
    return __builtin_bit_cast(unsigned short, (__bf16)v);
  }
  __device__ static __forceinline__ float to_f(unsigned short b) {
    return __uint_as_float((unsigned)b << 16);
  }
  // 16x16 MFMA with K = 32 (8 values = 4 packed words per lane) and K = 16
  __device__ static __forceinline__ v4f mfma32(v4u a, v4u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8, a),
                                                   __builtin_bit_cast(v8, b), c, 0, 0, 0);
  }
  __device__ static __forceinline__ v4f mfma16(v2u a, v2u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(v4s, a),
                                                     __builtin_bit_cast(v4s, b), c, 0, 0, 0);
  }
};

struct F16 {
  static constexpr int kPrec = DPK_PREC_F16;
  static constexpr bool kSplit = false;
  typedef __attribute__((ext_vector_type(2))) _Float16 v2;
  typedef __attribute__((ext_vector_type(4))) _Float16 v4;
  typedef __attribute__((ext_vector_type(8))) _Float16 v8;
  // two floats -> one packed pair, round to nearest even (v_cvt_pk_f16_f32)
  __device__ static __forceinline__ unsigned pack2(float a, float b) {
    return __builtin_bit_cast(unsigned, __builtin_convertvector(v2f{a, b}, v2));
  }
  __device__ static __forceinline__ float lo(unsigned u) {
    return (float)__builtin_bit_cast(v2, u)[0];
  }
  __device__ static __forceinline__ float hi(unsigned u) {
    return (float)__builtin_bit_cast(v2, u)[1];
  }
  __device__ static __forceinline__ unsigned short from_f(float v) {
    return __builtin_bit_cast(unsigned short, (_Float16)v);
  }
  __device__ static __forceinline__ float to_f(unsigned short b) {
    return (float)__builtin_bit_cast(_Float16, b);
  }
  __device__ static __forceinline__ v4f mfma32(v4u a, v4u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(v8, a),
                                                  __builtin_bit_cast(v8, b), c, 0, 0, 0);
  }
  __device__ static __forceinline__ v4f mfma16(v2u a, v2u b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x16f16(__builtin_bit_cast(v4, a),
                                                 __builtin_bit_cast(v4, b), c, 0, 0, 0);
  }
};

// FP16 weights with every data operand split into two FP16 values, hi =
// fp16(x) and lo = fp16(x - hi): a product with a weight takes two MFMAs
// (W hi + W lo), a product of two data operands three (hi hi + lo hi +
// hi lo); x is then carried to ~22 significant bits (absolute 2^-24 of an
// O(1) operand where lo is subnormal).
struct F16x2 : F16 {
  static constexpr int kPrec = DPK_PREC_F16X2;
  static constexpr bool kSplit = true;
};

// x - f32(h) for the low / high FP16 half h of a packed pair, one
// v_fma_mix_f32 each (FP16 source selected by op_sel, FP32 x, exact: the
// difference of a float and its nearest FP16 is representable).  Written
// out because the compiler turns fma(h, -1, x) into a subtraction, which
// costs a separate conversion of h.
__device__ __forceinline__ float sub_f16_lo(float x, unsigned pair) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(pair), "v"(x));
  return r;
}
__device__ __forceinline__ float sub_f16_hi(float x, unsigned pair) {
  float r;
  asm("v_fma_mix_f32 %0, %1, -1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=v"(r)
      : "v"(pair), "v"(x));
  return r;
}

// Storage tags of 16-bit tensors exchanged with the caller (GEMM operands).
struct bf16s {
  unsigned short bits;
  using fmt = Bf16;
};
struct f16s {
  unsigned short bits;
  using fmt = F16;
};

template <class T>
struct is_h16 {
  static constexpr bool value = false;
};
template <>
struct is_h16<bf16s> {
  static constexpr bool value = true;
};
template <>
struct is_h16<f16s> {
  static constexpr bool value = true;
};

// Power of two p with |x| p in [2^13, 2^14) for the largest |x| of a set
// (1 for 0): 16-bit operands scaled by p sit far above FP16's subnormal
// range (6.1e-5) and below its maximum; multiplying the FP32 result by 1/p
// undoes it exactly.
__device__ __forceinline__ float pow2_scale(float amax) {
  if (!(amax > 0.0f)) return 1.0f;
  int e = 0;
  (void)frexpf(amax, &e);  // amax = m 2^e, m in [0.5, 1)
  const int k = 14 - e;    // capped so that p stays a finite FP32
  return ldexpf(1.0f, k > 100 ? 100 : k);
}

}  // namespace h16
}  // namespace dpk
