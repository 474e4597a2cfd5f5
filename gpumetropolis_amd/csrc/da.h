// da.h -- per-iteration kernels of the stepwise sampler: plain Metropolis,
// two-stage delayed acceptance with a surrogate first stage, synchronous
// Differential Evolution and Langevin (MALA) steps.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vm.h"

namespace gpum {

// Read through a device pointer by one thread per chain.
struct DaArgs {
  uint32_t n_chains;
  uint32_t n_params;
  uint32_t n_iter;
  uint32_t seed;
  uint32_t stream_id;
  uint32_t iter_offset;
  // 0: diagonal sd; 1: Differential Evolution against the current
  // population, the pair excluding the proposing chain (the synchronous
  // path B, cpu_native.rs:400-506); 2: Cholesky factor per chain.
  uint32_t mode;
  double gamma, de_noise;  // mode 1
  const double* temperatures;
  const double* proposal_sd;
  const double* proposal_l;
  const uint32_t* prior;
  uint32_t n_prior;
  const double* prior_consts;
  // chain state (FP64)
  double* theta;       // n_chains x np
  double* prop;        // n_chains x np
  double* cur;         // exact log-posterior of theta
  double* cur_s;       // surrogate log-posterior of theta
  double* prop_prior;  // prior of prop
  double* prop_s;      // surrogate log-posterior of prop
  double* accepts;
  double* passes;      // proposals that reached the exact stage
  // exact stage: compacted list of proposals and their slice partials
  int32_t* pos;        // slot of the chain's proposal, or -1
  uint32_t* list;      // slot -> chain
  uint32_t* count;     // active slots
  const double* partial;     // slices x n_chains
  const uint32_t* n_slices;  // slices of the last exact-stage launch
  // surrogate first stage
  uint32_t use_surrogate;
  uint32_t n_in;
  const double* norm_mean;   // n_in
  const double* norm_w;      // n_in x n_in: x = W (theta - mean)
  double t_mean, t_sd;       // loglik = t_mean + t_sd * network output
  uint16_t* Xb;              // network input rows (bf16, 16 per row)
  const float* s_raw;        // network output per row
  // Langevin proposals (mode 3, all FP64): gradients of the log-posterior
  // at the states, the data term's value and gradient at the points just
  // evaluated (ll_prop, gy), the prior's (v_prior, g_prior), and the
  // proposals' means.
  double* gx;
  double* gy;
  double* mean_x;
  double* ll_prop;
  double* v_prior;
  double* g_prior;
  // outputs
  float* draws_it;       // iteration-major staging: (t * np + j) * C + c
  float* draws;          // reference layout t + n_iter (c + C j)
  double* trace_s;       // n_iter x C surrogate log-posterior of prop
  double* trace_plp;     // n_iter x C exact log-posterior (NaN: not run)
  uint8_t* trace_stage;  // n_iter x C: 0 stage-1 reject, 1 stage-2 reject, 2 accept
  double* trace_prop;    // n_iter x C x np proposals
  double* trace_ll;      // n_iter x C data term of the proposal (NaN: not run)
};

// Langevin steps (mode 3), one thread per chain: the log-posterior and
// gradient at the states from the data and prior terms in ll_prop / gy and
// v_prior / g_prior; the proposals; the decisions from the same terms at
// the proposals.
hipError_t launch_mala_init(const DaArgs* A, uint32_t n_chains, hipStream_t s);
hipError_t launch_mala_propose(const DaArgs* A, uint32_t n_chains, uint32_t t,
                               hipStream_t s);
hipError_t launch_mala_accept(const DaArgs* A, uint32_t n_chains, uint32_t t,
                              hipStream_t s);

hipError_t launch_da_init(const DaArgs* A, uint32_t n_chains, hipStream_t s);
hipError_t launch_da_init_finish(const DaArgs* A, uint32_t n_chains,
                                 hipStream_t s);
hipError_t launch_da_propose(const DaArgs* A, uint32_t n_chains, uint32_t t,
                             hipStream_t s);
hipError_t launch_da_stage1(const DaArgs* A, uint32_t t, hipStream_t s);
// draws (reference layout) from draws_it, once after the iterations.
hipError_t launch_da_draws(const float* draws_it, float* draws, uint32_t n_iter,
                           uint32_t rows, hipStream_t s);
hipError_t launch_da_stage2(const DaArgs* A, uint32_t n_chains, uint32_t t,
                            hipStream_t s);
// Network input rows of arbitrary points (surrogate predictions).
hipError_t launch_da_features(const double* points, uint32_t n, uint32_t n_in,
                              const double* mean, const double* w,
                              uint16_t* Xb, hipStream_t s);

}  // namespace gpum
