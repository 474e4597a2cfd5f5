// eval64.h -- FP64 evaluation of reference programs at many parameter points:
// the per-observation log-likelihood matrix (cpu_native.rs:371-391) and the
// gradient of the summed data term by reverse-mode differentiation
// (cpu_native.rs:148-279, 306-329).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#if !defined(__HIPCC_RTC__)
#include <vector>
#endif

namespace gpum {

// One reference instruction with the positions of its operands: the index
// of the instruction that produced each (every instruction produces one
// value).  l / r are unused by leaves, r by unary operations.  `slot`: where
// the gradient's forward pass keeps this instruction's value for the
// backward pass (kNoSlot when the backward pass never reads it; leaves are
// re-read instead of kept).
struct TapeOp {
  uint32_t op;
  uint32_t arg;
  uint32_t l;
  uint32_t r;
  uint32_t slot;
};
constexpr uint32_t kNoSlot = 0xFFFFFFFFu;

// Observations per (point, slice) work item of the gradient kernel: 256
// threads, 32 observations each.  It depends on nothing but the kernel, so
// a point's gradient has the same summation order in any batch.
constexpr uint32_t kGradSliceObs = 8192;
// Work items of the compiled pointwise kernel: an observation's
// parameter-free values are computed once for kPointGroup points, over
// slices of kPointwiseSliceObs observations (values do not depend on it).
constexpr uint32_t kPointGroup = 16;
constexpr uint32_t kPointwiseSliceObs = 1024;
// Longest program the gradient kernel's tape holds.
constexpr uint32_t kMaxTapeOps = 1024;

// Gradient work items (point, slice of kGradSliceObs observations), T =
// n_slices x n_points of them, taken slice-major (slice, then point) and
// cut into 8 runs of ceil(T / 8) items (the last run shorter), one per
// XCD: the points of a slice run together and share its rows in one XCD's
// L2 (a slice spans at most two XCDs), and the XCDs share the work
// evenly.  Item b is the (b / 8)-th of run b mod 8 and runs on XCD b mod 8
// (the hardware deals workgroups to the XCDs round-robin, and the grid is
// a multiple of 8 or covers every item); items past the last are empty.
constexpr uint32_t kEvalXcds = 8;
__host__ __device__ inline uint64_t grad64_items(uint32_t n_slices,
                                                 uint32_t n_points) {
  const uint64_t t = (uint64_t)n_slices * n_points;
  return kEvalXcds * ((t + kEvalXcds - 1) / kEvalXcds);
}

struct Eval64Args {
  const TapeOp* prog;
  uint32_t n_ops;
  uint32_t n_params;
  const double* consts;
  const double* data;    // column-major: column c at data[c * n_obs]
  uint32_t n_obs;
  uint32_t n_points;
  const double* points;  // n_points x n_params, point-major
  double* out;           // pointwise: n_points x n_obs; gradient: partials
  uint32_t n_slices;     // gradient: observation slices per point
  uint32_t depth;        // peak stack depth of the program
  uint32_t n_slots;      // gradient: values kept for the backward pass
};

#if !defined(__HIPCC_RTC__)
// The program with each instruction's operand positions and kept-value
// slots (host; the program must be valid, validate_program).  *n_slots
// (may be NULL): slots used.
std::vector<TapeOp> make_tape(const uint32_t* code, size_t n_ops,
                              uint32_t* n_slots = nullptr);
#endif

#if !defined(__HIPCC_RTC__)
// The program's kernel of jit kind `kind` (JIT_POINTWISE64 / JIT_GRAD64)
// compiled per program, or nullptr when compiled FP64 evaluation is
// switched off (gpum_hip_set_eval_jit) or unavailable in this process: the
// caller then runs the interpreter, which gives the same values.
hipFunction_t eval64_compiled_kernel(const uint32_t* code, size_t code_len,
                                     const double* consts, int kind);

// out[p * n_obs + i] = program at point p, observation i.
hipError_t launch_pointwise64(const Eval64Args& a, uint32_t grid,
                              hipStream_t st);
// partials[(p * n_slices + s) * (n_params + 1) + k]: d/d param k of the
// sum over slice s's observations at point p (k < n_params), and the sum
// itself (k = n_params).
hipError_t launch_grad64(const Eval64Args& a, uint32_t grid, hipStream_t st);
// out[p * n_params + k] = gradient summed over slices in slice order;
// value[p] (when not NULL) = the summed value.
hipError_t launch_grad64_finish(const double* partials, uint32_t n_points,
                                uint32_t n_params, uint32_t n_slices,
                                double* out, double* value, hipStream_t st);
#endif

}  // namespace gpum
