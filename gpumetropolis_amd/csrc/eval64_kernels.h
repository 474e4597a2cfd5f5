// eval64_kernels.h -- the FP64 batch-evaluation kernel bodies, generic over
// the per-observation evaluator E:
//   E::Hoisted, E::prepare(a, P)
//                          what depends on the point only, computed once per
//                          work item (generated code: parameter-only
//                          subexpressions; the interpreter: nothing);
//   E::value(a, h, P, i)   the program at point P, observation i;
//   E::grad(a, h, P, i, g) the same value, d/d param k added into g[k];
//   E::kFastGrad, E::grad_fast(a, h, P, i, g)
//                          (when kFastGrad) grad without the zero-adjoint
//                          rule: equal bits unless a sum of g ends NaN;
//                          i < 2^29 (an index within a slice).
// The interpreter (InterpEval64) and code generated per program share them.
//
// A work item is (point, slice of kGradSliceObs observations): 256 threads,
// 32 observations each, the point fixed across a thread's loop.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eval64.h"
#include "eval64_device.h"
#include "vm.h"

namespace gpum {

// A copy of x whose every 32-bit word is the first active lane's: for a
// value the whole wave holds alike, so that the compiler keeps it in
// scalar registers.
template <class T>
__device__ __forceinline__ T wave_uniform(const T& x) {
  if constexpr (sizeof(T) % 4 != 0) {
    return x;
  } else {
    uint32_t w[sizeof(T) / 4];
    __builtin_memcpy(w, &x, sizeof(T));
#pragma unroll
    for (uint32_t k = 0; k < sizeof(T) / 4; ++k)
      w[k] = __builtin_amdgcn_readfirstlane(w[k]);
    T r;
    __builtin_memcpy(&r, w, sizeof(T));
    return r;
  }
}

// out[p * n_obs + i] = the program at point p, observation i.
template <class E>
__device__ __forceinline__ void pointwise64_body(const Eval64Args& a) {
  const uint32_t ns = (a.n_obs + kGradSliceObs - 1) / kGradSliceObs;
  const uint64_t total = (uint64_t)ns * a.n_points;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t p = (uint32_t)(b / ns), s = (uint32_t)(b % ns);
    const double* P = a.points + (size_t)p * a.n_params;
    double* out = a.out + (size_t)p * a.n_obs;
    const typename E::Hoisted h = wave_uniform(E::prepare(a, P));
    const uint32_t i0 = s * kGradSliceObs + threadIdx.x;
    for (uint32_t j = 0; j < kGradSliceObs / 256u; ++j) {
      const uint32_t i = i0 + j * 256u;
      if (i < a.n_obs) out[i] = E::value(a, h, P, i);
    }
  }
}

// The pointwise matrix for evaluators that split an observation's
// parameter-free values (E::ObsVals, E::obs(a, i)) from the rest
// (E::value_o(a, h, o, P), the same operations as E::value): a work item is
// (group of up to kPointGroup points, slice of kPointwiseSliceObs
// observations, four per thread: enough items to fill every CU); each observation's parameter-free values (data loads,
// log y, ...) are computed once and serve every point of the group, whose
// hoisted values sit in LDS.  Writes stay coalesced: for each point the
// workgroup's threads write consecutive observations.

template <class E>
__device__ __forceinline__ void pointwise64_grouped_body(const Eval64Args& a) {
  __shared__ typename E::Hoisted hs[kPointGroup];
  const uint32_t ns = (a.n_obs + kPointwiseSliceObs - 1) / kPointwiseSliceObs;
  const uint32_t ng = (a.n_points + kPointGroup - 1) / kPointGroup;
  const uint64_t total = (uint64_t)ns * ng;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint32_t g = (uint32_t)(b / ns), s = (uint32_t)(b % ns);
    const uint32_t p0 = g * kPointGroup;
    const uint32_t gc = min(kPointGroup, a.n_points - p0);
    __syncthreads();  // the previous item is done with hs
    if (threadIdx.x < gc)
      hs[threadIdx.x] = E::prepare(a, a.points + (size_t)(p0 + threadIdx.x) * a.n_params);
    __syncthreads();
    const uint32_t i0 = s * kPointwiseSliceObs + threadIdx.x;
    for (uint32_t j = 0; j < kPointwiseSliceObs / 256u; ++j) {
      const uint32_t i = i0 + j * 256u;
      if (i >= a.n_obs) break;
      const typename E::ObsVals o = E::obs(a, i);
      for (uint32_t p = 0; p < gc; ++p)
        a.out[(size_t)(p0 + p) * a.n_obs + i] =
            E::value_o(a, hs[p], o, a.points + (size_t)(p0 + p) * a.n_params);
    }
  }
}

// partials[(p * n_slices + s) * (n_params + 1) + k]: the gradient (k <
// n_params) and the value (k = n_params) summed over slice s at point p.
// Lanes, then waves (butterfly), then the four waves in order.
template <class E>
__device__ __forceinline__ void grad64_body(const Eval64Args& a) {
  __shared__ double red[4][kMaxParams + 1];
  const uint32_t ns = a.n_slices;
  // The items in slice-major runs, one run per XCD (grad64_items), the
  // mapping in 64 bits (a 32-bit item counter made the compiler spill the
  // observation loop's registers).
  const uint64_t t = (uint64_t)ns * a.n_points;
  const uint64_t run = (t + kEvalXcds - 1) / kEvalXcds;
  const uint64_t total = kEvalXcds * run;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (uint64_t b = blockIdx.x; b < total; b += gridDim.x) {
    const uint64_t idx = (b % kEvalXcds) * run + b / kEvalXcds;
    if (idx >= t) continue;  // an empty item: uniform over the workgroup
    const uint32_t s = (uint32_t)(idx / a.n_points);
    const uint32_t p = (uint32_t)(idx % a.n_points);
    const double* P = a.points + (size_t)p * a.n_params;
    double g[kMaxParams];
#pragma unroll
    for (int k = 0; k < kMaxParams; ++k) g[k] = 0.0;
    double val = 0.0;
    const typename E::Hoisted h = wave_uniform(E::prepare(a, P));
    const uint32_t i0 = s * kGradSliceObs + threadIdx.x;
    bool exact = true;
    if constexpr (E::kFastGrad) {
      // Without the zero-adjoint rule: the same bits unless a sum ends
      // NaN, and then the whole item again with the rule.  The data are
      // addressed from the slice's first row, so that a load is a scalar
      // column base plus a 32-bit offset.
      Eval64Args as = a;
      as.data = a.data + (size_t)s * kGradSliceObs;
      const uint32_t left = a.n_obs - s * kGradSliceObs;
      for (uint32_t j = 0; j < kGradSliceObs / 256u; ++j) {
        const uint32_t li = threadIdx.x + j * 256u;
        if (li < left) val += E::grad_fast(as, h, P, li, g);
      }
      bool nan = false;
#pragma unroll
      for (int k = 0; k < kMaxParams; ++k) nan = nan || isnan(g[k]);
      exact = __syncthreads_or(nan) != 0;
      if (exact) {
#pragma unroll
        for (int k = 0; k < kMaxParams; ++k) g[k] = 0.0;
        val = 0.0;
      }
    }
    if (exact) {
      for (uint32_t j = 0; j < kGradSliceObs / 256u; ++j) {
        const uint32_t i = i0 + j * 256u;
        if (i < a.n_obs) val += E::grad(a, h, P, i, g);
      }
    }
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

// The interpreter: MAXD bounds the stack depth, MAXS the values the
// gradient keeps.
template <int MAXD, int MAXS>
struct InterpEval64 {
  static constexpr bool kFastGrad = false;
  struct Hoisted {};
  static __device__ __forceinline__ Hoisted prepare(const Eval64Args&,
                                                    const double*) {
    return Hoisted{};
  }
  static __device__ __forceinline__ double value(const Eval64Args& a,
                                                 const Hoisted&,
                                                 const double* P, uint32_t i) {
    return eval_obs64<MAXD>(a, P, i);
  }
  static __device__ __forceinline__ double grad(const Eval64Args& a,
                                                const Hoisted&,
                                                const double* P, uint32_t i,
                                                double (&g)[kMaxParams]) {
    return grad_obs64<MAXD, MAXS>(a, P, i, g);
  }
};

}  // namespace gpum
