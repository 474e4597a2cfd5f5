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
#include "math64.h"
#include "vm.h"

#pragma clang fp contract(off)

namespace gpum {
namespace {

__device__ __forceinline__ double data_at(const Eval64Args& a, uint32_t col,
                                          uint32_t i) {
  return a.data[(size_t)col * a.n_obs + i];
}

// The stack machine, cpu_native.rs:80-107.  The stack pointer follows the
// program only, so it is the same in every lane.
template <int MAXD>
__device__ double eval_obs64(const Eval64Args& a, const double* P,
                             uint32_t i) {
  double st[MAXD];
  int sp = 0;
  for (uint32_t pc = 0; pc < a.n_ops; ++pc) {
    const TapeOp o = a.prog[pc];
    switch (o.op) {
      case OP_PUSH_CONST: st[sp++] = a.consts[o.arg]; break;
      case OP_PUSH_PARAM: st[sp++] = P[o.arg]; break;
      case OP_PUSH_DATA: st[sp++] = data_at(a, o.arg, i); break;
      case OP_ADD: st[sp - 2] = st[sp - 2] + st[sp - 1]; --sp; break;
      case OP_SUB: st[sp - 2] = st[sp - 2] - st[sp - 1]; --sp; break;
      case OP_MUL: st[sp - 2] = st[sp - 2] * st[sp - 1]; --sp; break;
      case OP_DIV: st[sp - 2] = st[sp - 2] / st[sp - 1]; --sp; break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
      case OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case OP_LGAMMA: st[sp - 1] = lgamma64(st[sp - 1]); break;
      default: st[sp - 2] = pow(st[sp - 2], st[sp - 1]); --sp; break;
    }
  }
  return st[0];
}

// One observation's value and gradient, cpu_native.rs:148-279.  The forward
// pass records every instruction's value on the tape.  Each value is the
// operand of exactly one later instruction, so its adjoint receives exactly
// one contribution; walking the program backwards, the pending adjoints form
// a stack (the operands' adjoints are pushed, the right operand's last, and
// the right operand is the instruction just before).  A zero adjoint
// propagates zeros, as the reference skips it.  d/d param k accumulates
// into g[k].
template <int MAXT>
__device__ void grad_obs64(const Eval64Args& a, const double* P, uint32_t i,
                           double (&g)[kMaxParams]) {
  double v[MAXT];
  for (uint32_t pc = 0; pc < a.n_ops; ++pc) {
    const TapeOp o = a.prog[pc];
    double x;
    switch (o.op) {
      case OP_PUSH_CONST: x = a.consts[o.arg]; break;
      case OP_PUSH_PARAM: x = P[o.arg]; break;
      case OP_PUSH_DATA: x = data_at(a, o.arg, i); break;
      case OP_ADD: x = v[o.l] + v[o.r]; break;
      case OP_SUB: x = v[o.l] - v[o.r]; break;
      case OP_MUL: x = v[o.l] * v[o.r]; break;
      case OP_DIV: x = v[o.l] / v[o.r]; break;
      case OP_NEG: x = -v[o.l]; break;
      case OP_EXP: x = exp(v[o.l]); break;
      case OP_LOG: x = log(v[o.l]); break;
      case OP_SQRT: x = sqrt(v[o.l]); break;
      case OP_LGAMMA: x = lgamma64(v[o.l]); break;
      default: x = pow(v[o.l], v[o.r]); break;
    }
    v[pc] = x;
  }
  double adj[kMaxStack];
  int sp = 0;
  adj[sp++] = 1.0;
  for (int pc = (int)a.n_ops - 1; pc >= 0; --pc) {
    const TapeOp o = a.prog[pc];
    const double ga = adj[--sp];
    const bool live = ga != 0.0;
    double dl = 0.0, dr = 0.0;
    switch (o.op) {
      case OP_PUSH_CONST:
      case OP_PUSH_DATA:
        break;
      case OP_PUSH_PARAM:
        if (live) {
#pragma unroll
          for (int k = 0; k < kMaxParams; ++k)
            if ((uint32_t)k == o.arg) g[k] += ga;
        }
        break;
      case OP_ADD: dl = ga; dr = ga; break;
      case OP_SUB: dl = ga; dr = -ga; break;
      case OP_MUL: dl = ga * v[o.r]; dr = ga * v[o.l]; break;
      case OP_DIV: dl = ga / v[o.r]; dr = -(ga * v[pc] / v[o.r]); break;
      case OP_NEG: dl = -ga; break;
      case OP_EXP: dl = ga * v[pc]; break;
      case OP_LOG: dl = ga / v[o.l]; break;
      case OP_SQRT: dl = ga * 0.5 / v[pc]; break;
      case OP_LGAMMA: dl = ga * digamma64(v[o.l]); break;
      default: {
        const double base = v[o.l], ex = v[o.r];
        dl = ga * ex * pow(base, ex - 1.0);
        dr = base > 0.0 ? ga * v[pc] * log(base) : 0.0;
        break;
      }
    }
    if (o.op >= OP_ADD) {
      adj[sp++] = live ? dl : 0.0;
      if (o.op <= OP_DIV || o.op == OP_POW) adj[sp++] = live ? dr : 0.0;
    }
  }
}

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

template <int MAXT>
__global__ __launch_bounds__(256) void grad64_kernel(Eval64Args a) {
  __shared__ double red[4][kMaxParams];
  const uint32_t ns = a.n_slices;
  const uint64_t total = (uint64_t)ns * a.n_points;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t p = (uint32_t)(b / ns), s = (uint32_t)(b % ns);
    const double* P = a.points + (size_t)p * a.n_params;
    double g[kMaxParams];
#pragma unroll
    for (int k = 0; k < kMaxParams; ++k) g[k] = 0.0;
    const uint32_t i0 = s * kGradSliceObs + threadIdx.x;
    for (uint32_t j = 0; j < kGradSliceObs / 256u; ++j) {
      const uint32_t i = i0 + j * 256u;
      if (i < a.n_obs) grad_obs64<MAXT>(a, P, i, g);
    }
    // Wave sums (butterfly), then the four waves in order.
#pragma unroll
    for (int k = 0; k < kMaxParams; ++k) {
      double x = g[k];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
      if (lane == 0) red[w][k] = x;
    }
    __syncthreads();
    if (threadIdx.x < a.n_params) {
      const int k = threadIdx.x;
      const double t = ((red[0][k] + red[1][k]) + red[2][k]) + red[3][k];
      a.out[((size_t)p * ns + s) * a.n_params + k] = t;
    }
    __syncthreads();
  }
}

__global__ void grad64_finish_kernel(const double* part, uint32_t n_points,
                                     uint32_t np, uint32_t ns, double* out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_points * np) return;
  const uint32_t p = t / np, k = t % np;
  double s = 0.0;
  for (uint32_t j = 0; j < ns; ++j) s += part[((size_t)p * ns + j) * np + k];
  out[t] = s;
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
  if (a.n_ops <= 64)
    hipLaunchKernelGGL(grad64_kernel<64>, dim3(grid), dim3(256), 0, st, a);
  else if (a.n_ops <= 256)
    hipLaunchKernelGGL(grad64_kernel<256>, dim3(grid), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(grad64_kernel<kMaxTapeOps>, dim3(grid), dim3(256), 0,
                       st, a);
  return hipGetLastError();
}

hipError_t launch_grad64_finish(const double* partials, uint32_t n_points,
                                uint32_t n_params, uint32_t n_slices,
                                double* out, hipStream_t st) {
  const uint32_t n = n_points * n_params;
  hipLaunchKernelGGL(grad64_finish_kernel, dim3((n + 255) / 256), dim3(256), 0,
                     st, partials, n_points, n_params, n_slices, out);
  return hipGetLastError();
}

}  // namespace gpum
