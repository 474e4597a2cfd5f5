// chain_kernels.h -- the interpreter's kernels for gfx950, templates over
// the VM variant, and the variants the library instantiates (one group per
// chain_kernel_inst*.hip file, compiled in parallel).
//
// Replaces metropolis_block_kernel (src/rust/src/chain_kernel.rs:122-357).
// Layout of one workgroup: `cpb` chains, `wpc` waves per chain (at most 16
// waves).  Per iteration:
//   1. data pass: every wave evaluates its chain's VM program over its share
//      of each LDS data tile (tiles stream from HBM through LDS and are
//      shared by all chains of the workgroup, or stay resident when the
//      whole data set fits), accumulating FP32 chunk sums into an FP64
//      per-lane accumulator;
//   2. reduction: FP64 cross-lane butterfly in the wave, then the chain's
//      wave partials in LDS;
//   3. fused step on the chain's lead lane, all in FP64: finish iteration
//      t (prior + Metropolis decision log u < (plp - cur) / T, record the
//      draw) and prepare t + 1 (Philox uniforms, Box-Muller, diagonal /
//      Cholesky / DE proposal, hoisted parameter program -> uniform table,
//      prior of the proposal).
// The first data pass evaluates the initial state (cur = log-posterior of
// init), exactly as the reference does before its loop.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "chain_body.h"

namespace gpum {

template <int NS, int OPT, int CW, bool STREAM>
__global__ void __launch_bounds__(1024)
chain_kernel(const ChainArgs* __restrict__ Ap, ChainHot H) {
  chain_body<VmEval<NS, OPT, CW>, STREAM>(Ap, H, blockIdx.x * H.cpb);
}

template <int NS, int OPT, int CW>
__global__ void __launch_bounds__(1024)
chain_kernel_batch(const BatchQueue* __restrict__ q,
                   const ChainArgs* __restrict__ args,
                   const ChainHot* __restrict__ hots) {
  batch_body<VmEval<NS, OPT, CW>>(q, args, hots);
}

template <int NS, int OPT, int CW, bool STREAM>
__global__ void __launch_bounds__(1024)
points_kernel(const PointsArgs* __restrict__ P, ChainHot H) {
  points_body<VmEval<NS, OPT, CW>, STREAM>(P, H);
}

// The affine-program interpreter (VmAffineEval).
template <int OPT, int CW, bool STREAM>
__global__ void __launch_bounds__(1024)
chain_kernel_affine(const ChainArgs* __restrict__ Ap, ChainHot H) {
  chain_body<VmAffineEval<OPT, CW>, STREAM>(Ap, H, blockIdx.x * H.cpb);
}

template <int OPT, int CW>
__global__ void __launch_bounds__(1024)
chain_kernel_batch_affine(const BatchQueue* __restrict__ q,
                          const ChainArgs* __restrict__ args,
                          const ChainHot* __restrict__ hots) {
  batch_body<VmAffineEval<OPT, CW>>(q, args, hots);
}

template <int OPT, int CW, bool STREAM>
__global__ void __launch_bounds__(1024)
points_kernel_affine(const PointsArgs* __restrict__ P, ChainHot H) {
  points_body<VmAffineEval<OPT, CW>, STREAM>(P, H);
}

// Affine programs run with 16 observations per lane (variant (0, 16, cw):
// no stack, the only programs the host gives 16 observations per lane
// without one); one, two or four chains per wave.
#define GPUM_AFFINE_VARIANTS(X) \
  X(16, 1)                      \
  X(16, 2)                      \
  X(16, 4)

#define GPUM_AFFINE_KERNELS(PRE, OPT, CW)                                     \
  PRE template __global__ void chain_kernel_affine<OPT, CW, false>(          \
      const ChainArgs*, ChainHot);                                            \
  PRE template __global__ void chain_kernel_affine<OPT, CW, true>(           \
      const ChainArgs*, ChainHot);                                            \
  PRE template __global__ void points_kernel_affine<OPT, CW, false>(         \
      const PointsArgs*, ChainHot);                                           \
  PRE template __global__ void points_kernel_affine<OPT, CW, true>(          \
      const PointsArgs*, ChainHot);                                           \
  PRE template __global__ void chain_kernel_batch_affine<OPT, CW>(           \
      const BatchQueue*, const ChainArgs*, const ChainHot*);
#define GPUM_AFFINE_INSTANTIATE(OPT, CW) GPUM_AFFINE_KERNELS(, OPT, CW)
#define GPUM_AFFINE_DECLARE(OPT, CW) GPUM_AFFINE_KERNELS(extern, OPT, CW)

// The general interpreter's variants: (spill slots, observations per lane,
// chains per wave).  Two chains per wave only for programs without a stack;
// 8 or 4 observations per lane for those that are not affine (variant
// (0, 16, cw) runs on the affine kernels above).
#define GPUM_VARIANTS(X) \
  X(0, 8, 1)             \
  X(0, 8, 2)             \
  X(0, 4, 1)             \
  X(0, 4, 2)             \
  X(1, 16, 1)            \
  X(3, 16, 1)            \
  X(7, 8, 1)             \
  X(15, 4, 1)            \
  X(31, 4, 1)

// Explicit instantiation (PRE empty) or its declaration (PRE extern) of
// one variant's kernels.
#define GPUM_KERNELS(PRE, NS, OPT, CW)                                        \
  PRE template __global__ void chain_kernel<NS, OPT, CW, false>(             \
      const ChainArgs*, ChainHot);                                            \
  PRE template __global__ void chain_kernel<NS, OPT, CW, true>(              \
      const ChainArgs*, ChainHot);                                            \
  PRE template __global__ void points_kernel<NS, OPT, CW, false>(            \
      const PointsArgs*, ChainHot);                                           \
  PRE template __global__ void points_kernel<NS, OPT, CW, true>(             \
      const PointsArgs*, ChainHot);                                           \
  PRE template __global__ void chain_kernel_batch<NS, OPT, CW>(              \
      const BatchQueue*, const ChainArgs*, const ChainHot*);
#define GPUM_INSTANTIATE(NS, OPT, CW) GPUM_KERNELS(, NS, OPT, CW)
#define GPUM_DECLARE(NS, OPT, CW) GPUM_KERNELS(extern, NS, OPT, CW)

}  // namespace gpum
