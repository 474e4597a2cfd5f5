// batch_kernels.hip -- the between-batch step of a many-model run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "batch.h"

namespace gpum {

__global__ void batch_end_kernel(const BatchModelInfo* __restrict__ models,
                                 const uint32_t* __restrict__ chain_model,
                                 uint32_t n_chains, uint32_t t0, uint32_t bsize,
                                 uint32_t b, int adapt) {
  // Separate roundings, as the host orchestrator computes these moments.
#pragma clang fp contract(off)
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_chains) return;
  const BatchModelInfo M = models[chain_model[g]];
  const uint32_t c = g - M.chain_off;
  const uint32_t np = M.n_params, C = M.n_chains;
  const size_t ld = M.draws_ld;
  if (M.accept_hist) M.accept_hist[(size_t)(b - 1) * C + c] = M.accept[c];
  if (adapt) {
    double* w = M.welford + (size_t)c * (1 + 2 * np);
    double* mean = w + 1;
    double* m2 = w + 1 + np;
    for (uint32_t t = t0; t < t0 + bsize; ++t) {
      w[0] += 1.0;
      for (uint32_t j = 0; j < np; ++j) {
        const double x = (double)M.draws[t + ld * (c + (size_t)C * j)];
        const double d1 = x - mean[j];
        mean[j] += d1 / w[0];
        m2[j] += d1 * (x - mean[j]);
      }
    }
    for (uint32_t j = 0; j < np; ++j) {
      // the seed covariance (initial sd^2) until two draws
      const double s0 = M.sd0[(size_t)c * np + j];
      const double var = w[0] >= 2.0 ? m2[j] / (w[0] - 1.0) : s0 * s0;
      M.sd[(size_t)c * np + j] = sqrt(fmax(var, 1e-12));
    }
  }
  const uint32_t last = t0 + bsize - 1;
  for (uint32_t j = 0; j < np; ++j)
    M.snapshot[(size_t)c * np + j] =
        (double)M.draws[last + ld * (c + (size_t)C * j)];
}

hipError_t launch_batch_end(const BatchModelInfo* models,
                            const uint32_t* chain_model, uint32_t n_chains,
                            uint32_t t0, uint32_t bsize, uint32_t b,
                            int adapt, hipStream_t s) {
  hipLaunchKernelGGL(batch_end_kernel, dim3((n_chains + 255) / 256), dim3(256),
                     0, s, models, chain_model, n_chains, t0, bsize, b, adapt);
  return hipGetLastError();
}

}  // namespace gpum
