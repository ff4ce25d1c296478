// Three-way BF16 split of FP32 values for matrix-core products (Henry, Tang &
// Heinecke, "Leveraging the bfloat16 artificial intelligence datatype for
// higher-precision computations", ARITH 2019): x = b0 + b1 + b2 exactly (bf16
// each, round-to-nearest-even), and
//   x*y ~= b0c0 + (b0c1 + b1c0) + (b0c2 + b1c1 + b2c0),
// dropping the terms below 2^-24 relative.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace xt {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf16_to_f(unsigned short h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ void split3(float x, unsigned short &h0, unsigned short &h1, unsigned short &h2)
{
  h0 = bf16_rne(x);
  const float r1 = x - bf16_to_f(h0);
  h1 = bf16_rne(r1);
  const float r2 = r1 - bf16_to_f(h1);
  h2 = bf16_rne(r2);
}

__device__ __forceinline__ f4 mfma_bf(const us8 &a, const us8 &b, f4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), c, 0, 0, 0);
}

}  // namespace xt
