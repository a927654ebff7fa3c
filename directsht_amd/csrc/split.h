// fp16 hi/lo splitting for matrix-core (MFMA) operands: x = hi + lo with
// hi = fp16(x) (round to nearest even) and lo = fp16(x - hi), so that
// hi*hi + hi*lo + lo*hi reproduces an FP32 product to ~2^-22.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsht {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

typedef _Float16 half2v __attribute__((ext_vector_type(2)));
typedef float float2v __attribute__((ext_vector_type(2)));

// x = hi + lo for a pair of values.  The pair conversion is v_cvt_pk_f16_f32
// (round to nearest even); lo = fp16(x - hi) is written as fma(hi, -1, x),
// which the compiler emits as v_fma_mix{lo,hi}_f16 reading the very hi
// register the MFMA reads (x - hi is exact in FP32, so lo is correctly
// rounded).  Plain code rather than inline assembly lets the compiler track
// the VALU-write -> MFMA-read hazards and interleave the split with MFMAs.
__device__ inline void split_pair(float a, float b, half2v& hi, half2v& lo) {
  const float2v v = {a, b};
  hi = __builtin_convertvector(v, half2v);
  lo[0] = (_Float16)__builtin_fmaf((float)hi[0], -1.0f, a);
  lo[1] = (_Float16)__builtin_fmaf((float)hi[1], -1.0f, b);
}

// hi/lo of 8 values (one MFMA operand slice)
__device__ inline void split8_f16(const float (&x)[8], half8& ahi, half8& alo) {
  u32x4 h, l;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    half2v ph, pl;
    split_pair(x[2 * j], x[2 * j + 1], ph, pl);
    h[j] = __builtin_bit_cast(uint32_t, ph);
    l[j] = __builtin_bit_cast(uint32_t, pl);
  }
  ahi = __builtin_bit_cast(half8, h);
  alo = __builtin_bit_cast(half8, l);
}

// hi/lo of two values (a, b), each packed as (a | b << 16), for operands that
// are staged through LDS as separate halves (no register dependency on an
// MFMA).  Inline assembly here: with the halves stored one by one the compiler
// would otherwise emit convert/subtract sequences instead of v_fma_mix.
__device__ inline void split2_f16(float a, float b, uint32_t& hi, uint32_t& lo) {
  asm("v_cvt_pk_f16_f32 %0, %2, %3\n\t"
      "v_fma_mixlo_f16 %1, -%0, 1.0, %2 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %1, -%0, 1.0, %3 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
      : "=&v"(hi), "=&v"(lo)
      : "v"(a), "v"(b));
}

__device__ inline _Float16 half_lo(uint32_t u) { return __builtin_bit_cast(_Float16, (uint16_t)(u & 0xffffu)); }
__device__ inline _Float16 half_hi(uint32_t u) { return __builtin_bit_cast(_Float16, (uint16_t)(u >> 16)); }

}  // namespace dsht
