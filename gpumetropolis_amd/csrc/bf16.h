// bf16.h -- FP32 -> bf16 conversion used by the surrogate's kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

// Round to nearest even; a NaN stays a quiet NaN.
__host__ __device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u;
  __builtin_memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)0x7fc0u;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

}  // namespace gpum
