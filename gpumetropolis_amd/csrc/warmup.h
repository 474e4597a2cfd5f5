// warmup.h -- the reference's adaptive warmup run on the device.
//
// .am_orchestrate_warmup (R/adaptive.R:101-254) calls the chain kernel once
// per batch of ~50 iterations and adapts on the host between batches:
// per-chain Welford moments (their diagonal sets the per-dimension scale),
// a Welford accumulator pooled over all chains (its regularised Cholesky
// factor, times each chain's Robbins-Monro scalar, is the next batch's
// proposal L), a Robbins-Monro step per chain toward the target
// acceptance, a restart of every accumulator at mid-warmup, and an
// optional early freeze.  Here the between-batch step is three small
// kernels after each batch launch, and the whole warmup (every batch and
// every step) is one HIP graph: nothing returns to the host until the
// warmup ends.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

// Device state of one warmup.  Per-chain arrays are chain-major.
struct WarmupDev {
  uint32_t n_chains, np, n_batches, reset_batch;
  int32_t auto_stop;
  double target;     // target acceptance
  double scale0;     // Robbins-Monro scalar after a restart
  const uint32_t* sizes;  // iterations of each batch
  const double* gain;     // Robbins-Monro gain of each batch, b^(-2/3)
  // the batch just run
  const float* draws;   // t + T (c + C j), T = this batch's length
  const float* accept;  // C
  // chain state handed to the next batch
  double* init;    // C x np: the last f32 draw of each chain
  double* sd;      // C x np
  double* scale;   // C
  double* prev_scale;
  double* L;       // C x np x np, row-major per chain
  // per-chain Welford moments (diagonal; the off-diagonal terms are never
  // read) and the batch statistics the pooled accumulator merges
  double* w_n;     // C
  double* w_mean;  // C x np
  double* w_m2;    // C x np
  double* w_sig;   // C x np: each chain's diagonal seed covariance
  double* b_mean;  // C x np: the batch's mean of each chain
  double* b_m2;    // C x np x np: the batch's deviation products
  // pooled accumulator
  double* p_n;     // 1
  double* p_mean;  // np
  double* p_m2;    // np x np
  double* p_sig;   // np: diagonal seed covariance
  double* L_pool;  // np x np
  double* diag_now;  // np: pooled variance at the restart
  // bookkeeping: has_L, frozen_at (0: running), batches run, iterations
  // used, freeze streak
  uint32_t* ctl;
  float* acc_hist;  // n_batches x C
};

enum WarmupCtl : uint32_t {
  kWarmHasL = 0,
  kWarmFrozenAt = 1,
  kWarmBatches = 2,
  kWarmUsed = 3,
  kWarmStreak = 4,
  kWarmCtlWords = 8
};

// Batch b (1-based) has run: per-chain moments, batch statistics,
// Robbins-Monro step, next start state and scales, acceptance history.
hipError_t launch_warmup_chains(const WarmupDev& w, uint32_t b,
                                uint32_t batch_len, hipStream_t s);
// The pooled accumulator absorbs batch b chain by chain; at the restart
// batch every accumulator restarts; the regularised Cholesky factor; the
// early-freeze test.
hipError_t launch_warmup_pool(const WarmupDev& w, uint32_t b,
                              uint32_t batch_len, hipStream_t s);
// Per chain: the restart of the chain's moments and scalar, then
// L_c = scale_c * L_pool.
hipError_t launch_warmup_post(const WarmupDev& w, uint32_t b,
                              hipStream_t s);
// Zero the device state of a warmup and seed it: each chain's seed
// covariance is its initial sd^2, the pooled one chain 0's
// (R/adaptive.R:119-121, 140-144); every scalar starts at scale0.
hipError_t launch_warmup_init(const WarmupDev& w, const double* sd0,
                              hipStream_t s);

}  // namespace gpum
