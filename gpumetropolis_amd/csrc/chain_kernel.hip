// chain_kernel.hip -- host launchers of the interpreter's kernels
// (chain_kernels.h); the kernels are instantiated in chain_kernel_inst*.hip.
// Variant (0, 16, cw) -- affine programs -- runs on the affine interpreter.
#include <hip/hip_runtime.h>

#include "chain_kernels.h"

namespace gpum {

GPUM_VARIANTS(GPUM_DECLARE)
GPUM_AFFINE_VARIANTS(GPUM_AFFINE_DECLARE)

hipError_t launch_chain_kernel(const ChainArgs* dev_args, const ChainHot& h,
                               int ns, int opt, int cw, size_t lds_bytes,
                               hipStream_t s) {
  dim3 grid((h.n_chains + h.cpb - 1) / h.cpb);
  dim3 block(h.cpb / cw * h.wpc * 64);
  const bool stream = h.stream_mode != 0;
#define GPUM_AFFINE_CASE(OPT_, CW_)                                          \
  if (ns == 0 && opt == OPT_ && cw == CW_) {                                 \
    if (stream)                                                              \
      hipLaunchKernelGGL((chain_kernel_affine<OPT_, CW_, true>), grid, block, \
                         lds_bytes, s, dev_args, h);                         \
    else                                                                     \
      hipLaunchKernelGGL((chain_kernel_affine<OPT_, CW_, false>), grid,      \
                         block, lds_bytes, s, dev_args, h);                  \
    return hipGetLastError();                                                \
  }
  GPUM_AFFINE_VARIANTS(GPUM_AFFINE_CASE)
#undef GPUM_AFFINE_CASE
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
#define GPUM_AFFINE_CASE(OPT_, CW_)                                          \
  if (ns == 0 && opt == OPT_ && cw == CW_) {                                 \
    if (stream)                                                              \
      hipLaunchKernelGGL((points_kernel_affine<OPT_, CW_, true>), grid,      \
                         block, lds_bytes, s, p, h);                         \
    else                                                                     \
      hipLaunchKernelGGL((points_kernel_affine<OPT_, CW_, false>), grid,     \
                         block, lds_bytes, s, p, h);                         \
    return hipGetLastError();                                                \
  }
  GPUM_AFFINE_VARIANTS(GPUM_AFFINE_CASE)
#undef GPUM_AFFINE_CASE
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

hipError_t launch_chain_kernel_batch(const BatchQueue* q, uint32_t n_wg,
                                     const ChainArgs* args,
                                     const ChainHot* hots, int ns, int opt,
                                     int cw, size_t lds_bytes, hipStream_t s) {
#define GPUM_AFFINE_CASE(OPT_, CW_)                                          \
  if (ns == 0 && opt == OPT_ && cw == CW_) {                                 \
    hipLaunchKernelGGL((chain_kernel_batch_affine<OPT_, CW_>), dim3(n_wg),   \
                       dim3(1024), lds_bytes, s, q, args, hots);             \
    return hipGetLastError();                                                \
  }
  GPUM_AFFINE_VARIANTS(GPUM_AFFINE_CASE)
#undef GPUM_AFFINE_CASE
#define GPUM_CASE(NS_, OPT_, CW_)                                            \
  if (ns == NS_ && opt == OPT_ && cw == CW_) {                               \
    hipLaunchKernelGGL((chain_kernel_batch<NS_, OPT_, CW_>), dim3(n_wg),     \
                       dim3(1024), lds_bytes, s, q, args, hots);             \
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
#define GPUM_AFFINE_ATTR(OPT_, CW_)                                              \
  e = hipFuncSetAttribute((const void*)chain_kernel_affine<OPT_, CW_, false>,    \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)bytes);                                           \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)chain_kernel_affine<OPT_, CW_, true>,     \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds));                      \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)points_kernel_affine<OPT_, CW_, false>,   \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kPointsExtraLds));                       \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)points_kernel_affine<OPT_, CW_, true>,    \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds - kPointsExtraLds));    \
  if (e != hipSuccess) return e;                                                 \
  e = hipFuncSetAttribute((const void*)chain_kernel_batch_affine<OPT_, CW_>,     \
                          hipFuncAttributeMaxDynamicSharedMemorySize,            \
                          (int)(bytes - kStreamStaticLds - kBatchArgsLds));      \
  if (e != hipSuccess) return e;
  GPUM_AFFINE_VARIANTS(GPUM_AFFINE_ATTR)
#undef GPUM_AFFINE_ATTR
  return e;
}

}  // namespace gpum
