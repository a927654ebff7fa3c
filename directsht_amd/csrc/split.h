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

// x = hi + lo for 8 values: hi = fp16(x) (round to nearest even, packed in
// pairs), lo = fp16(x - hi) computed by a mixed-precision FMA that reads the
// very hi register the MFMA will read (x - hi is exact in FP32, so lo is
// correctly rounded).  One asm block keeps the compiler from deriving hi and
// lo from differently rounded copies of x; the trailing s_nop covers the
// VALU-write -> MFMA-read dependency, which the compiler's hazard tracking
// does not see through inline assembly.
__device__ inline void split8_f16(const float (&x)[8], half8& ahi, half8& alo) {
  uint32_t h0, h1, h2, h3, l0, l1, l2, l3;
  asm("v_cvt_pk_f16_f32 %0, %8, %9\n\t"
      "v_cvt_pk_f16_f32 %1, %10, %11\n\t"
      "v_cvt_pk_f16_f32 %2, %12, %13\n\t"
      "v_cvt_pk_f16_f32 %3, %14, %15\n\t"
      "v_fma_mixlo_f16 %4, -%0, 1.0, %8 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %4, -%0, 1.0, %9 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %5, -%1, 1.0, %10 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %5, -%1, 1.0, %11 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %6, -%2, 1.0, %12 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %6, -%2, 1.0, %13 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %7, -%3, 1.0, %14 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %7, -%3, 1.0, %15 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "s_nop 4"
      : "=&v"(h0), "=&v"(h1), "=&v"(h2), "=&v"(h3), "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3)
      : "v"(x[0]), "v"(x[1]), "v"(x[2]), "v"(x[3]), "v"(x[4]), "v"(x[5]), "v"(x[6]), "v"(x[7]));
  ahi = __builtin_bit_cast(half8, (u32x4){h0, h1, h2, h3});
  alo = __builtin_bit_cast(half8, (u32x4){l0, l1, l2, l3});
}

// hi/lo of two values (a, b), each packed as (a | b << 16); for operands that
// are staged through LDS (no VALU -> MFMA register dependency).
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
