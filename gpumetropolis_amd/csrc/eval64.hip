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
// The interpreter's program is wave-uniform: each instruction is read with
// a scalar load, every lane evaluates it for its own observation.  Data
// reads are coalesced (consecutive lanes, consecutive observations of a
// column).  The kernel bodies are in eval64_kernels.h, shared with the
// kernels compiled per program.
#include "eval64.h"
#include "eval64_device.h"
#include "eval64_kernels.h"
#include "math64.h"
#include "vm.h"

namespace gpum {
namespace {

template <int MAXD>
__global__ __launch_bounds__(256) void pointwise64_kernel(Eval64Args a) {
  pointwise64_body<InterpEval64<MAXD, 1>>(a);
}

// At least two waves per SIMD (at most 256 VGPRs): with one the FP64
// arithmetic's latency is exposed, with four the stacks spill to scratch
// (logistic-16, 33 points x 1e6 rows: 50 ms, 28.7 ms, 176 ms).
template <int MAXD, int MAXS>
__global__ __launch_bounds__(256, 2) void grad64_kernel(Eval64Args a) {
  grad64_body<InterpEval64<MAXD, MAXS>>(a);
}

__global__ void grad64_finish_kernel(const double* part, uint32_t n_points,
                                     uint32_t np, uint32_t ns, double* out,
                                     double* value) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_points * (np + 1)) return;
  const uint32_t p = t / (np + 1), k = t % (np + 1);
  const double* q = part + (size_t)p * ns * (np + 1) + k;
  const size_t ld = np + 1;
  // Slice order; sixteen slices' partials are loaded before they are added,
  // so that the loads overlap.
  double s = 0.0;
  uint32_t j = 0;
  for (; j + 16 <= ns; j += 16) {
    double v[16];
#pragma unroll
    for (uint32_t u = 0; u < 16; ++u) v[u] = q[(j + u) * ld];
#pragma unroll
    for (uint32_t u = 0; u < 16; ++u) s += v[u];
  }
  for (; j < ns; ++j) s += q[j * ld];
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
