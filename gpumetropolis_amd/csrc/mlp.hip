// mlp.hip -- kernels of the surrogate network (16-256-256-1, SiLU).
//
// The two wide layers and every weight gradient are products of bf16
// matrices on the gfx950 matrix cores (v_mfma_f32_16x16x32_bf16, or
// 16x16x16 where K is 16; FP32 accumulation); the 256 -> 1 head, its
// backward, the loss and Adam are VALU kernels.  Activations that feed a matrix product are stored in bf16;
// pre-activations, the last hidden layer and all gradients stay FP32.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include "bf16.h"
#include "mlp.h"
#include "philox.h"

namespace gpum {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef short v8s __attribute__((ext_vector_type(8)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float sigmoidf_(float x) {
  return 1.0f / (1.0f + __expf(-x));
}
__device__ __forceinline__ float silu(float x) { return x * sigmoidf_(x); }
// d/dx x sigma(x) = sigma(x) (1 + x (1 - sigma(x)))
__device__ __forceinline__ float silu_grad(float x) {
  const float sg = sigmoidf_(x);
  return sg * (1.0f + x * (1.0f - sg));
}

// One wave per 16 x 16 output tile, four waves per workgroup along N.
// Operand fragments (lane l, r = l % 16, q = l / 16): for
// v_mfma_f32_16x16x32_bf16 A holds A(r, 8q + i) and B holds B(8q + i, r),
// i = 0..7; for v_mfma_f32_16x16x16_bf16 A(r, 4q + i) and B(4q + i, r),
// i = 0..3.  A and B take the same k for each (lane, i), so the product
// sums every k whichever order the instruction adds them in.  The result
// holds D(4q + i, r), i = 0..3.
__global__ void __launch_bounds__(256) gemm_bf16_kernel(GemmArgs g) {
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 16;
  const int n0 = blockIdx.x * 64 + wave * 16;
  const int r = lane & 15, q = lane >> 4;
  const uint16_t* __restrict__ A = g.A;
  const uint16_t* __restrict__ B = g.B;
  v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
  const size_t arow = (size_t)(m0 + r) * g.sam;
  const size_t bcol = (size_t)(n0 + r) * g.sbn;
  if (g.K % 128 == 0) {
    // Four 32-deep slabs per step, their loads issued together: the
    // operands are small and cache-resident, and a wave's loop is bound by
    // load latency, not by the matrix unit.
    for (int k0 = 0; k0 < g.K; k0 += 128) {
      v8s a[4], b[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const size_t k = (size_t)(k0 + 32 * u + 8 * q + i);
          a[u][i] = (short)A[arow + k * g.sak];
          b[u][i] = (short)B[k * g.sbk + bcol];
        }
#pragma unroll
      for (int u = 0; u < 4; ++u)
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
            __builtin_bit_cast(v8bf, a[u]), __builtin_bit_cast(v8bf, b[u]), acc,
            0, 0, 0);
    }
  } else if (g.K % 32 == 0) {
    for (int k0 = 0; k0 < g.K; k0 += 32) {
      v8s a, b;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const size_t k = (size_t)(k0 + 8 * q + i);
        a[i] = (short)A[arow + k * g.sak];
        b[i] = (short)B[k * g.sbk + bcol];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
          __builtin_bit_cast(v8bf, a), __builtin_bit_cast(v8bf, b), acc, 0, 0,
          0);
    }
  } else {
    for (int k0 = 0; k0 < g.K; k0 += 16) {
      v4s a, b;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const size_t k = (size_t)(k0 + 4 * q + i);
        a[i] = (short)A[arow + k * g.sak];
        b[i] = (short)B[k * g.sbk + bcol];
      }
      acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
  }
  const int n = n0 + r;
  float csum = 0.0f;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int m = m0 + 4 * q + i;
    const size_t o = (size_t)m * g.ld + n;
    float v = acc[i];
    if (g.bias) v += g.bias[n];
    if (g.dact) v *= silu_grad(g.dact[o]);
    if (g.out) g.out[o] = v;
    if (g.out_b) g.out_b[o] = f2bf(v);
    if (g.act || g.act_b) {
      const float h = silu(v);
      if (g.act) g.act[o] = h;
      if (g.act_b) g.act_b[o] = f2bf(h);
    }
    csum += v;
  }
  if (g.colsum) {
    // rows 4q..4q+3 of column n in this lane; add the four q groups
    csum += __shfl_xor(csum, 16, 64);
    csum += __shfl_xor(csum, 32, 64);
    if (q == 0) g.colsum[(size_t)(m0 / 16) * g.ld + n] = csum;
  }
}

hipError_t launch_gemm_bf16(const GemmArgs& g, hipStream_t s) {
  if (g.M <= 0 || g.N <= 0 || g.K <= 0 || g.M % 16 || g.N % 64 || g.K % 16)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(gemm_bf16_kernel, dim3(g.N / 64, g.M / 16), dim3(256), 0,
                     s, g);
  return hipGetLastError();
}

// One wave per row.
__global__ void __launch_bounds__(256) mlp_head_kernel(const float* __restrict__ H2,
                                                       const float* __restrict__ P,
                                                       float* __restrict__ s,
                                                       int M) {
  const int lane = threadIdx.x & 63;
  const int m = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (m >= M) return;
  const float* h = H2 + (size_t)m * kMlpHid;
  const float* w3 = P + kOffW3;
  float acc = 0.0f;
#pragma unroll
  for (int k = 0; k < kMlpHid / 64; ++k) acc += h[lane + 64 * k] * w3[lane + 64 * k];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if (lane == 0) s[m] = acc + P[kOffB3];
}

hipError_t launch_mlp_head(const float* H2, const float* params, float* s,
                           int M, hipStream_t st) {
  if (M <= 0) return hipSuccess;
  hipLaunchKernelGGL(mlp_head_kernel, dim3((M + 3) / 4), dim3(256), 0, st, H2,
                     params, s, M);
  return hipGetLastError();
}

__device__ __forceinline__ float head_g(const float* s, const float* t, int m,
                                        int valid) {
  return m < valid ? 2.0f * (s[m] - t[m]) / (float)valid : 0.0f;
}

// grid (kMlpHid / 64, M / 64); 64 columns x 4 row groups per workgroup.
__global__ void __launch_bounds__(256) mlp_head_bwd_kernel(
    const float* __restrict__ s, const float* __restrict__ t,
    const float* __restrict__ H2, const float* __restrict__ A2,
    const float* __restrict__ P, uint16_t* __restrict__ D2b,
    float* __restrict__ part, int valid) {
  __shared__ float sh_w[4][64], sh_b[4][64];
  const int cl = threadIdx.x & 63, rg = threadIdx.x >> 6;
  const int j = blockIdx.x * 64 + cl;
  const int mb = blockIdx.y * 64;
  const float w3 = P[kOffW3 + j];
  float aw = 0.0f, ab = 0.0f;
  for (int i = 0; i < 16; ++i) {
    const int m = mb + rg + 4 * i;
    const float gm = head_g(s, t, m, valid);
    const size_t o = (size_t)m * kMlpHid + j;
    const float d = gm * w3 * silu_grad(A2[o]);
    D2b[o] = f2bf(d);
    aw += gm * H2[o];
    ab += d;
  }
  sh_w[rg][cl] = aw;
  sh_b[rg][cl] = ab;
  __syncthreads();
  float* pr = part + (size_t)blockIdx.y * kPartStride;
  if (rg == 0) {
    pr[j] = ((sh_w[0][cl] + sh_w[1][cl]) + sh_w[2][cl]) + sh_w[3][cl];
    pr[256 + j] = ((sh_b[0][cl] + sh_b[1][cl]) + sh_b[2][cl]) + sh_b[3][cl];
  }
  if (blockIdx.x == 0 && rg == 1) {
    const int m = mb + cl;
    float gm = head_g(s, t, m, valid);
    float se = m < valid ? (s[m] - t[m]) * (s[m] - t[m]) : 0.0f;
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      gm += __shfl_xor(gm, off, 64);
      se += __shfl_xor(se, off, 64);
    }
    if (cl == 0) {
      pr[512] = gm;
      pr[513] = se;
    }
  }
}

hipError_t launch_mlp_head_bwd(const float* s, const float* t, const float* H2,
                               const float* A2, const float* params,
                               uint16_t* D2b, float* part, int M, int valid,
                               hipStream_t st) {
  if (M <= 0 || M % 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_head_bwd_kernel, dim3(kMlpHid / 64, M / 64), dim3(256),
                     0, st, s, t, H2, A2, params, D2b, part, valid);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) mlp_grad_finish_kernel(
    const float* __restrict__ part, const float* __restrict__ cs_b1, int M,
    float* __restrict__ grad, float* __restrict__ loss, MlpSched* sched,
    float* __restrict__ losses, MlpTrainConst tc) {
  const int j = threadIdx.x;
  float b1 = 0.0f;
  for (int r = 0; r < M / 16; ++r) b1 += cs_b1[(size_t)r * kMlpHid + j];
  float w3 = 0.0f, b2 = 0.0f;
  for (int r = 0; r < M / 64; ++r) {
    w3 += part[(size_t)r * kPartStride + j];
    b2 += part[(size_t)r * kPartStride + 256 + j];
  }
  grad[kOffB1 + j] = b1;
  grad[kOffB2 + j] = b2;
  grad[kOffW3 + j] = w3;
  if (j == 0) {
    float b3 = 0.0f, se = 0.0f;
    for (int r = 0; r < M / 64; ++r) {
      b3 += part[(size_t)r * kPartStride + 512];
      se += part[(size_t)r * kPartStride + 513];
    }
    grad[kOffB3] = b3;
    if (loss) *loss = se;
    if (sched) {
      const uint32_t k = sched->k, t = sched->t + 1u;
      if (losses) losses[k] = se;
      // cosine decay from lr to lr_final over the call's steps
      const double frac =
          tc.steps > 1 ? (double)k / (double)(tc.steps - 1) : 1.0;
      const double lr = tc.lr_final + 0.5 * (tc.lr - tc.lr_final) *
                                          (1.0 + cos(3.141592653589793 * frac));
      const double bc1 = 1.0 - pow((double)tc.beta1, (double)t);
      const double bc2 = 1.0 - pow((double)tc.beta2, (double)t);
      sched->lr_t = (float)(lr * sqrt(bc2) / bc1);
      sched->t = t;
      sched->k = k + 1u;
    }
  }
}

hipError_t launch_mlp_grad_finish(const float* part, const float* colsum_b1,
                                  int M, float* grad, float* loss,
                                  hipStream_t st, MlpSched* sched,
                                  float* losses, const MlpTrainConst* tc) {
  MlpTrainConst c;
  memset(&c, 0, sizeof c);
  if (tc) c = *tc;
  hipLaunchKernelGGL(mlp_grad_finish_kernel, dim3(1), dim3(kMlpHid), 0, st, part,
                     colsum_b1, M, grad, loss, sched, losses, c);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) adam_kernel(
    float* __restrict__ P, const float* __restrict__ G, float* __restrict__ Mm,
    float* __restrict__ V, uint16_t* __restrict__ W1b,
    uint16_t* __restrict__ W2b, float lr_t, float b1, float b2, float eps,
    const MlpSched* __restrict__ sched) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kMlpParams) return;
  if (sched) lr_t = sched->lr_t;
  const float g = G[i];
  const float m = b1 * Mm[i] + (1.0f - b1) * g;
  const float v = b2 * V[i] + (1.0f - b2) * g * g;
  Mm[i] = m;
  V[i] = v;
  const float p = P[i] - lr_t * m / (sqrtf(v) + eps);
  P[i] = p;
  if (i >= kOffW1 && i < kOffB1) W1b[i - kOffW1] = f2bf(p);
  if (i >= kOffW2 && i < kOffB2) W2b[i - kOffW2] = f2bf(p);
}

hipError_t launch_adam(float* params, const float* grad, float* m, float* v,
                       uint16_t* W1b, uint16_t* W2b, float lr_t, float beta1,
                       float beta2, float eps, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0,
                     st, params, grad, m, v, W1b, W2b, lr_t, beta1, beta2, eps,
                     (const MlpSched*)nullptr);
  return hipGetLastError();
}

hipError_t launch_adam_sched(float* params, const float* grad, float* m,
                             float* v, uint16_t* W1b, uint16_t* W2b,
                             const MlpSched* sched, float beta1, float beta2,
                             float eps, hipStream_t st) {
  hipLaunchKernelGGL(adam_kernel, dim3((kMlpParams + 255) / 256), dim3(256), 0,
                     st, params, grad, m, v, W1b, W2b, 0.0f, beta1, beta2, eps,
                     sched);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) mlp_mirror_kernel(const float* __restrict__ P,
                                                         uint16_t* __restrict__ W1b,
                                                         uint16_t* __restrict__ W2b) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < kMlpIn * kMlpHid) W1b[i] = f2bf(P[kOffW1 + i]);
  if (i < kMlpHid * kMlpHid) W2b[i] = f2bf(P[kOffW2 + i]);
}

hipError_t launch_mlp_mirror(const float* params, uint16_t* W1b, uint16_t* W2b,
                             hipStream_t st) {
  hipLaunchKernelGGL(mlp_mirror_kernel, dim3(kMlpHid * kMlpHid / 256), dim3(256),
                     0, st, params, W1b, W2b);
  return hipGetLastError();
}

__global__ void __launch_bounds__(256) mlp_gather_kernel(
    const uint16_t* __restrict__ Xtr, const float* __restrict__ Ttr,
    uint32_t n_rows, uint32_t seed, uint32_t step, uint16_t* __restrict__ Xb,
    float* __restrict__ tb, int M, int valid,
    const MlpSched* __restrict__ sched) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= M) return;
  if (sched) step = sched->k;
  if (b >= valid || n_rows == 0) {
    for (int k = 0; k < kMlpIn; ++k) Xb[(size_t)b * kMlpIn + k] = 0;
    tb[b] = 0.0f;
    return;
  }
  const U4 h = philox4x32_10(U4{step, (uint32_t)b, 0u, TAG_MLP}, seed, 0u);
  const uint32_t r = h.x % n_rows;
  for (int k = 0; k < kMlpIn; ++k)
    Xb[(size_t)b * kMlpIn + k] = Xtr[(size_t)r * kMlpIn + k];
  tb[b] = Ttr[r];
}

hipError_t launch_mlp_gather(const uint16_t* Xtr, const float* Ttr,
                             uint32_t n_rows, uint32_t seed, uint32_t step,
                             uint16_t* Xb, float* tb, int M, int valid,
                             hipStream_t st, const MlpSched* sched) {
  hipLaunchKernelGGL(mlp_gather_kernel, dim3((M + 255) / 256), dim3(256), 0, st,
                     Xtr, Ttr, n_rows, seed, step, Xb, tb, M, valid, sched);
  return hipGetLastError();
}

}  // namespace gpum
