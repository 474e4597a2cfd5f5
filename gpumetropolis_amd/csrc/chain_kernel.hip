// chain_kernel.hip -- the fused Metropolis chain kernel for gfx950.
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
chain_kernel_batch(const BatchWork* __restrict__ work,
                   const ChainArgs* __restrict__ args,
                   const ChainHot* __restrict__ hots, BatchStep step) {
  batch_body<VmEval<NS, OPT, CW>>(work, args, hots, step);
}

template <int NS, int OPT, int CW, bool STREAM>
__global__ void __launch_bounds__(1024)
points_kernel(const PointsArgs* __restrict__ P, ChainHot H) {
  points_body<VmEval<NS, OPT, CW>, STREAM>(P, H);
}

// The interpreter's variants: (spill slots, observations per lane, chains
// per wave).  Two chains per wave only for programs without a stack.
#define GPUM_VARIANTS(X) \
  X(0, 16, 1)            \
  X(0, 16, 2)            \
  X(1, 16, 1)            \
  X(3, 16, 1)            \
  X(7, 8, 1)             \
  X(15, 4, 1)            \
  X(31, 4, 1)

#define GPUM_INST(NS, OPT, CW)                                                \
  template __global__ void chain_kernel<NS, OPT, CW, false>(const ChainArgs*, \
                                                            ChainHot);        \
  template __global__ void chain_kernel<NS, OPT, CW, true>(const ChainArgs*,  \
                                                           ChainHot);         \
  template __global__ void points_kernel<NS, OPT, CW, false>(                 \
      const PointsArgs*, ChainHot);                                           \
  template __global__ void points_kernel<NS, OPT, CW, true>(const PointsArgs*, \
                                                            ChainHot);        \
  template __global__ void chain_kernel_batch<NS, OPT, CW>(                   \
      const BatchWork*, const ChainArgs*, const ChainHot*, BatchStep);
GPUM_VARIANTS(GPUM_INST)
#undef GPUM_INST

hipError_t launch_chain_kernel(const ChainArgs* dev_args, const ChainHot& h,
                               int ns, int opt, int cw, size_t lds_bytes,
                               hipStream_t s) {
  dim3 grid((h.n_chains + h.cpb - 1) / h.cpb);
  dim3 block(h.cpb / cw * h.wpc * 64);
  const bool stream = h.stream_mode != 0;
#define GPUM_CASE(NS_, OPT_, CW_)                                            \
  if (ns == NS_ && opt == OPT_ && cw == CW_) {                               \
    if (stream)                                                              \
      hipLaunchKernelGGL((chain_kernel<NS_, OPT_, CW_, true>), grid, block,  \
                         lds_bytes, s, dev_args, h);                         \
    else                                                                     \
      hipLaunchKernelGGL((chain_kernel<NS_, OPT_, CW_, false>), grid, block, \
                         lds_bytes, s, dev_args, h);                         \
    return hipGetLastError();                                                \
  }
  GPUM_VARIANTS(GPUM_CASE)
#undef GPUM_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_points_kernel(const PointsArgs* p, uint32_t n_workgroups,
                                const ChainHot& h, int ns, int opt, int cw,
                                size_t lds_bytes, hipStream_t s) {
  dim3 grid(n_workgroups);
  dim3 block(h.cpb / cw * 64);
  const bool stream = h.stream_mode != 0;
#define GPUM_CASE(NS_, OPT_, CW_)                                            \
  if (ns == NS_ && opt == OPT_ && cw == CW_) {                               \
    if (stream)                                                              \
      hipLaunchKernelGGL((points_kernel<NS_, OPT_, CW_, true>), grid, block, \
                         lds_bytes, s, p, h);                                \
    else                                                                     \
      hipLaunchKernelGGL((points_kernel<NS_, OPT_, CW_, false>), grid,       \
                         block, lds_bytes, s, p, h);                         \
    return hipGetLastError();                                                \
  }
  GPUM_VARIANTS(GPUM_CASE)
#undef GPUM_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_chain_kernel_batch(const BatchWork* work, uint32_t n_wg,
                                     const ChainArgs* args,
                                     const ChainHot* hots, BatchStep step,
                                     int ns, int opt, int cw,
                                     size_t lds_bytes, hipStream_t s) {
#define GPUM_CASE(NS_, OPT_, CW_)                                            \
  if (ns == NS_ && opt == OPT_ && cw == CW_) {                               \
    hipLaunchKernelGGL((chain_kernel_batch<NS_, OPT_, CW_>), dim3(n_wg),     \
                       dim3(1024), lds_bytes, s, work, args, hots, step);    \
    return hipGetLastError();                                                \
  }
  GPUM_VARIANTS(GPUM_CASE)
#undef GPUM_CASE
  return hipErrorInvalidValue;
}

hipError_t set_chain_kernel_lds_limit(size_t bytes) {
  hipError_t e = hipSuccess;
#define GPUM_ATTR(NS_, OPT_, CW_)                                                \
  e = hipFuncSetAttribute((const void*)chain_kernel<NS_, OPT_, CW_, false>,      \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)bytes);                                           \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)chain_kernel<NS_, OPT_, CW_, true>,       \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds));                      \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)points_kernel<NS_, OPT_, CW_, false>,     \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kPointsExtraLds));                       \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)points_kernel<NS_, OPT_, CW_, true>,      \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds - kPointsExtraLds));    \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)chain_kernel_batch<NS_, OPT_, CW_>,       \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds - kBatchArgsLds));      \
  if (e != hipSuccess) return e;
  GPUM_VARIANTS(GPUM_ATTR)
#undef GPUM_ATTR
  return e;
}

}  // namespace gpum
