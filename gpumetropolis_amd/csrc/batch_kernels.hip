// batch_kernels.hip -- the sync-word reset of the persistent batched kernel.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "batch.h"

namespace gpum {

__global__ void batch_sync_reset_kernel(uint32_t* __restrict__ sync,
                                        uint32_t n_words) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_words) sync[i] = 0u;
}

hipError_t launch_batch_sync_reset(uint32_t* sync, uint32_t n_words,
                                   hipStream_t s) {
  hipLaunchKernelGGL(batch_sync_reset_kernel, dim3((n_words + 255) / 256),
                     dim3(256), 0, s, sync, n_words);
  return hipGetLastError();
}

}  // namespace gpum
