// batch.h -- between-batch step of a many-model run (one thread per chain).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

// Per-model view of the flat per-chain buffers of a batched run.
struct BatchModelInfo {
  uint32_t chain_off;  // first global chain of the model
  uint32_t n_chains;
  uint32_t n_params;
  uint32_t draws_ld;     // iterations of the whole run
  const float* draws;    // reference layout t + draws_ld (c + C j)
  double* snapshot;      // C x np: start states / DE population
  double* sd;            // C x np proposal scale
  const double* sd0;     // C x np initial proposal scale (Welford seed)
  double* welford;       // C x (1 + 2 np): count, mean, M2
  const float* accept;   // C: acceptance of the batch just run
  float* accept_hist;    // n_batches x C
};

// End of batch b (1-based) covering draw rows t0 .. t0 + bsize - 1:
// the population snapshot becomes the last draws (the reference's host
// hands back f32 draws, R/adaptive.R:500-502), the batch acceptance is
// recorded, and with `adapt` the per-chain Welford moments absorb the
// batch's draws and sd = sqrt(max(diag cov, 1e-12)) (R/adaptive.R:487-497).
hipError_t launch_batch_end(const BatchModelInfo* models,
                            const uint32_t* chain_model, uint32_t n_chains,
                            uint32_t t0, uint32_t bsize, uint32_t b,
                            int adapt, hipStream_t s);

}  // namespace gpum
