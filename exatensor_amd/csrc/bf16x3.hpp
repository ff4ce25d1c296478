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

// Tile-blocked split planes: blocks of RB rows x KB k-values laid out as a
// matrix kernel's LDS stage holds them, so that one wave-wide 16-byte DMA
// moves a whole block (RB * KB * 2 bytes, 1 KB for 16 x 32) from memory
// into LDS. Lane l of a block holds row l / (KB / 8) at 16-byte slot
// l % (KB / 8).
template <int RB, int KB>
struct B3Layout {
  static constexpr int kChunks = KB / 8;                  // 16-byte chunks per block row
  // chunk c of row r sits at c ^ ((r >> kShift) & (kChunks - 1)); for the
  // 16-row blocks, row bits 1-2 are what leaves the fragment reads free of
  // bank conflicts (SQ_LDS_BANK_CONFLICT 0, against 4 cycles per ds_read_b128
  // with bits 0-1, 2-3 or 3-4, and 6% more matrix throughput)
  static constexpr int kShift = 1;
  __host__ __device__ static int pos(int r, int c) { return c ^ ((r >> kShift) & (kChunks - 1)); }
};

// Wave-wide 16-byte DMA from global memory straight into LDS: lane i's 16
// bytes land at lds_wave_base + 16 i.
__device__ __forceinline__ void lds_dma16(const void *g, void *lds_wave_base)
{
  __builtin_amdgcn_global_load_lds(g, lds_wave_base, 16, 0, 0);
}

// s_waitcnt with only the vector-memory count (expcnt and lgkmcnt at their
// no-wait maxima), gfx9 encoding.
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
  __builtin_amdgcn_s_waitcnt((N & 15) | (7 << 4) | (15 << 8) | ((N >> 4) << 14));
}

// s_waitcnt vmcnt(0) with the LDS and export counters left free.
__device__ __forceinline__ void wait_all_vmem() { wait_vmcnt<0>(); }

// Barrier without the release/acquire fences of __syncthreads (whose
// workgroup-scope release waits for every outstanding load, including the
// DMA of later stages); the callers order LDS with explicit vmcnt waits. The
// empty asm statements keep the compiler from moving memory accesses across.
__device__ __forceinline__ void stage_barrier()
{
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace xt
