// mlp.h -- the surrogate network's kernels: a 16-256-256-1 SiLU MLP with
// BF16 MFMA matrix products (FP32 accumulation), its backward pass and Adam.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

constexpr int kMlpIn = 16;    // input features (zero-padded below 16)
constexpr int kMlpHid = 256;  // hidden width of both layers

// Offsets of the flat FP32 parameter vector (gradients use the same).
constexpr int kOffW1 = 0;                          // kMlpIn x kMlpHid
constexpr int kOffB1 = kOffW1 + kMlpIn * kMlpHid;  // kMlpHid
constexpr int kOffW2 = kOffB1 + kMlpHid;           // kMlpHid x kMlpHid
constexpr int kOffB2 = kOffW2 + kMlpHid * kMlpHid; // kMlpHid
constexpr int kOffW3 = kOffB2 + kMlpHid;           // kMlpHid
constexpr int kOffB3 = kOffW3 + kMlpHid;           // 1
constexpr int kMlpParams = kOffB3 + 1;

// Training schedule on the device, so a captured graph of training steps
// needs no per-step arguments: k counts this call's steps (the minibatch
// draw and the learning-rate schedule), t Adam's steps over the network's
// life; lr_t is the step's bias-corrected rate.
struct MlpSched {
  uint32_t k, t;
  float lr_t;
  uint32_t pad;
};

// The constants of a training call.
struct MlpTrainConst {
  uint32_t steps;
  float beta1, beta2, eps;
  double lr, lr_final;
};

// Head-backward partials of one 64-row block: [0, 256) sum g*H2 (w3),
// [256, 512) sum D2 (b2), 512 sum g (b3), 513 squared error.
constexpr int kPartStride = 516;

// C = op(A) op(B) over bf16 operands with FP32 accumulation, M and K
// multiples of 16, N a multiple of 64 (K a multiple of 32 runs the 32-deep
// matrix instruction).  A(m, k) = A[m*sam + k*sak],
// B(k, n) = B[k*sbk + n*sbn].  Epilogue per element, in this order: add
// bias[n]; multiply by silu'(dact[m*ld + n]); store the value (out, FP32;
// out_b, bf16); store silu(value) (act, FP32; act_b, bf16); colsum gets
// per-16-row column sums of the value, row block r at colsum[r*ld + n].
struct GemmArgs {
  const uint16_t* A;
  int sam, sak;
  const uint16_t* B;
  int sbk, sbn;
  int M, N, K;
  const float* bias;
  const float* dact;
  float* out;
  uint16_t* out_b;
  float* act;
  uint16_t* act_b;
  float* colsum;
  int ld;
};

hipError_t launch_gemm_bf16(const GemmArgs& g, hipStream_t s);

// s[m] = b3 + sum_j H2[m, j] w3[j] for m < M (rows of 256).
hipError_t launch_mlp_head(const float* H2, const float* params, float* s,
                           int M, hipStream_t st);

// Backward of the head and the mean-squared-error loss over the first
// `valid` of M rows (rows past it get zero gradient):
//   g[m] = 2 (s[m] - t[m]) / valid,
//   D2[m, j] = g[m] w3[j] silu'(A2[m, j])  (bf16 out),
// with per-64-row partial sums of g H2 (for w3), D2 (for b2), g and the
// squared error written to `part` (layout in mlp.hip).
hipError_t launch_mlp_head_bwd(const float* s, const float* t, const float* H2,
                               const float* A2, const float* params,
                               uint16_t* D2b, float* part, int M, int valid,
                               hipStream_t st);

// Sums the partials into the gradient vector (b1, b2, w3, b3) and `loss`
// (the sum of squared errors over the valid rows).  With `sched`: the
// squared error also goes to losses[sched->k], and the step's rate
// (cosine decay from lr to lr_final over the call's steps, Adam's bias
// corrections at step t + 1) to sched->lr_t; then k and t advance.
hipError_t launch_mlp_grad_finish(const float* part, const float* colsum_b1,
                                  int M, float* grad, float* loss,
                                  hipStream_t st, MlpSched* sched = nullptr,
                                  float* losses = nullptr,
                                  const MlpTrainConst* tc = nullptr);

// One Adam step over the flat parameters; refreshes the bf16 copies of W1
// and W2.  lr_t carries the bias corrections.
hipError_t launch_adam(float* params, const float* grad, float* m, float* v,
                       uint16_t* W1b, uint16_t* W2b, float lr_t, float beta1,
                       float beta2, float eps, hipStream_t st);
// The same step with the rate read from sched->lr_t.
hipError_t launch_adam_sched(float* params, const float* grad, float* m,
                             float* v, uint16_t* W1b, uint16_t* W2b,
                             const MlpSched* sched, float beta1, float beta2,
                             float eps, hipStream_t st);

// Rebuild the bf16 copies of W1 and W2 from the FP32 parameters.
hipError_t launch_mlp_mirror(const float* params, uint16_t* W1b, uint16_t* W2b,
                             hipStream_t st);

// Minibatch gather: row b of the batch is training row
// philox(step, b, 0, TAG_MLP) mod n_rows; rows past `valid` are zero.
// With `sched`, step is sched->k.
hipError_t launch_mlp_gather(const uint16_t* Xtr, const float* Ttr,
                             uint32_t n_rows, uint32_t seed, uint32_t step,
                             uint16_t* Xb, float* tb, int M, int valid,
                             hipStream_t st, const MlpSched* sched = nullptr);

}  // namespace gpum
