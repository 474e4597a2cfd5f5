// eval64.hip -- FP64 kernels evaluating a reference program at many
// parameter points: the per-observation values (cpu_native.rs:371-391) and
// the gradient of the summed data term by reverse-mode differentiation of
// the bytecode (cpu_native.rs:148-279, 306-329).
//
// Both read the data as FP64 columns and evaluate every operation in FP64
// with the reference's expressions, one rounding per operation (no fused
// multiply-add), so values agree with the reference's CPU path up to the
// last-place differences of exp / log / pow between math libraries.
//
// The program is wave-uniform: each instruction is read with a scalar load,
// every lane evaluates it for its own observation.  Data reads are coalesced
// (consecutive lanes, consecutive observations of a column).
#include "eval64.h"
#include "eval64_device.h"
#include "math64.h"
#include "vm.h"

#pragma clang fp contract(off)

namespace gpum {
namespace {

template <int MAXD>
__global__ __launch_bounds__(256) void pointwise64_kernel(Eval64Args a) {
  const uint32_t nb = (a.n_obs + 255u) / 256u;
  const uint64_t total = (uint64_t)nb * a.n_points;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t p = (uint32_t)(b / nb);
    const uint32_t i = (uint32_t)(b % nb) * 256u + threadIdx.x;
    const uint32_t ic = i < a.n_obs ? i : a.n_obs - 1;
    const double x = eval_obs64<MAXD>(a, a.points + (size_t)p * a.n_params, ic);
    if (i < a.n_obs) a.out[(size_t)p * a.n_obs + i] = x;
  }
}

// At least two waves per SIMD (at most 256 VGPRs): with one the FP64
// arithmetic's latency is exposed, with four the stacks spill to scratch
// (logistic-16, 33 points x 1e6 rows: 50 ms, 28.7 ms, 176 ms).
template <int MAXD, int MAXS>
__global__ __launch_bounds__(256, 2) void grad64_kernel(Eval64Args a) {
  __shared__ double red[4][kMaxParams + 1];
  const uint32_t ns = a.n_slices;
  const uint64_t total = (uint64_t)ns * a.n_points;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t p = (uint32_t)(b / ns), s = (uint32_t)(b % ns);
    const double* P = a.points + (size_t)p * a.n_params;
    double g[kMaxParams];
#pragma unroll
    for (int k = 0; k < kMaxParams; ++k) g[k] = 0.0;
    double val = 0.0;
    const uint32_t i0 = s * kGradSliceObs + threadIdx.x;
    for (uint32_t j = 0; j < kGradSliceObs / 256u; ++j) {
      const uint32_t i = i0 + j * 256u;
      if (i < a.n_obs) val += grad_obs64<MAXD, MAXS>(a, P, i, g);
    }
    // Wave sums (butterfly), then the four waves in order; slot kMaxParams
    // carries the value.
#pragma unroll
    for (int k = 0; k <= kMaxParams; ++k) {
      double x = k < kMaxParams ? g[k] : val;
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
      if (lane == 0) red[w][k] = x;
    }
    __syncthreads();
    const uint32_t np = a.n_params;
    if (threadIdx.x <= np) {
      const int k = threadIdx.x < np ? (int)threadIdx.x : kMaxParams;
      const double t = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
      a.out[((size_t)p * ns + s) * (np + 1) + threadIdx.x] = t;
    }
    __syncthreads();
  }
}

__global__ void grad64_finish_kernel(const double* part, uint32_t n_points,
                                     uint32_t np, uint32_t ns, double* out,
                                     double* value) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_points * (np + 1)) return;
  const uint32_t p = t / (np + 1), k = t % (np + 1);
  double s = 0.0;
  for (uint32_t j = 0; j < ns; ++j)
    s += part[((size_t)p * ns + j) * (np + 1) + k];
  if (k < np)
    out[(size_t)p * np + k] = s;
  else if (value)
    value[p] = s;
}

}  // namespace

hipError_t launch_pointwise64(const Eval64Args& a, uint32_t grid,
                              hipStream_t st) {
  if (a.depth <= 8)
    hipLaunchKernelGGL(pointwise64_kernel<8>, dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(pointwise64_kernel<kMaxStack>, dim3(grid), dim3(256), 0,
                       st, a);
  return hipGetLastError();
}

hipError_t launch_grad64(const Eval64Args& a, uint32_t grid, hipStream_t st) {
  // Shallow stacks and few kept values stay in registers.
  if (a.depth <= 8 && a.n_slots <= 16)
    hipLaunchKernelGGL((grad64_kernel<8, 16>), dim3(grid), dim3(256), 0, st, a);
  else if (a.depth <= 8 && a.n_slots <= 64)
    hipLaunchKernelGGL((grad64_kernel<8, 64>), dim3(grid), dim3(256), 0, st, a);
  else if (a.n_slots <= 64)
    hipLaunchKernelGGL((grad64_kernel<kMaxStack, 64>), dim3(grid), dim3(256), 0,
                       st, a);
  else
    hipLaunchKernelGGL((grad64_kernel<kMaxStack, kMaxTapeOps>), dim3(grid),
                       dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_grad64_finish(const double* partials, uint32_t n_points,
                                uint32_t n_params, uint32_t n_slices,
                                double* out, double* value, hipStream_t st) {
  const uint32_t n = n_points * (n_params + 1);
  hipLaunchKernelGGL(grad64_finish_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     st, partials, n_points, n_params, n_slices, out, value);
  return hipGetLastError();
}

}  // namespace gpum
