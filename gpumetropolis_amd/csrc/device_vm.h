// device_vm.h -- the wavefront-uniform interpreter of the per-observation
// VM (vm.h) for gfx950.
//
// One wave evaluates the program of one chain over a chunk of 64 * OPT
// observations: lane l holds observations [4*l + 256*g + j] (g < OPT/4,
// j < 4) of the chunk, read from the LDS data tile with 128-bit loads.  The
// instruction stream is read with scalar loads (the program pointer and pc
// are uniform), the operands of every case are statically indexed
// registers: T[OPT] is the cached top of stack, S[NS][OPT] the spill slots.
// Uniform operands (hoisted parameter values and constants) come from the
// chain's table in LDS as broadcast reads.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "vm.h"

namespace gpum {

// ln(2) in FP32 for natural logs from the hardware log2.
#define GPUM_LN2F 0.69314718055994531f
#define GPUM_LOG2EF 1.4426950408889634f

__device__ __forceinline__ float vm_exp(float x) {
  return __builtin_amdgcn_exp2f(x * GPUM_LOG2EF);
}
__device__ __forceinline__ float vm_log(float x) {
  return __builtin_amdgcn_logf(x) * GPUM_LN2F;
}
__device__ __forceinline__ float vm_rcp(float x) {
  return __builtin_amdgcn_rcpf(x);
}
__device__ __forceinline__ float vm_sqrt(float x) {
  return __builtin_amdgcn_sqrtf(x);
}

// pow with the sign rules of C pow for a negative base and an integral
// exponent; a negative base with a non-integral exponent gives NaN.
__device__ __forceinline__ float vm_pow(float b, float e) {
  float ab = fabsf(b);
  float r = __builtin_amdgcn_exp2f(e * __builtin_amdgcn_logf(ab));
  float et = truncf(e);
  bool e_int = et == e;
  // odd integral exponent: e/2 is not integral (exact for |e| < 2^24)
  bool e_odd = e_int && (truncf(e * 0.5f) != e * 0.5f);
  r = (b < 0.0f && e_odd) ? -r : r;
  r = (b < 0.0f && !e_int) ? __builtin_nanf("") : r;
  r = (e == 0.0f || b == 1.0f) ? 1.0f : r;
  return r;
}

// FP32 log-gamma of a positive argument, for the data-dependent lgamma
// opcode.
//   x >= 8:     Stirling's series with three correction terms (truncation
//               error below 3e-10).
//   0 < x < 8:  shift to z = x - k in [2, 3), k = floor(x) - 2:
//                 k = -2, -1:  lgamma(x) = lgamma(z) - ln(x (x + 1)) or - ln x
//                 k = 1..5:    lgamma(x) = lgamma(z) + ln((x-1) ... (x-k))
//               (every factor x - j is exact), and lgamma(z) = (z - 2) q(z)
//               with q a degree-7 polynomial in z - 2.5 (coefficients from
//               csrc/lgamma_fit.py: least squares against a 50-digit
//               log-gamma, rounded to FP32).
// One log serves both branches.  The relative error is about 2e-7 of
// max(|lgamma(x)|, 1), where Stirling after shifting up to 8 lost ~3e-6.
__device__ __forceinline__ float vm_lgamma_pos(float x) {
  const bool big = x >= 8.0f;
  const float k = floorf(x) - 2.0f;
  const float z = x - k;
  float pd = 1.0f;
#pragma unroll
  for (int j = 1; j <= 5; ++j)
    pd = (float)j <= k ? pd * (x - (float)j) : pd;
  const float pu = x < 1.0f ? x * (x + 1.0f) : x;
  const float p = k < 0.0f ? pu : pd;
  const float lg = vm_log(big ? x : p);
  const float t = z - 2.5f;
  float q = 1.149600794e-04f;
  q = fmaf(q, t, -3.346131707e-04f);
  q = fmaf(q, t, 9.595991578e-04f);
  q = fmaf(q, t, -3.092080355e-03f);
  q = fmaf(q, t, 1.087710634e-02f);
  q = fmaf(q, t, -4.480599985e-02f);
  q = fmaf(q, t, 2.675817907e-01f);
  q = fmaf(q, t, 5.693657398e-01f);
  const float poly = (z - 2.0f) * q;
  const float small = k < 0.0f ? poly - lg : poly + lg;
  const float inv = vm_rcp(x);
  const float inv2 = inv * inv;
  const float series =
      inv * (0.083333333333333333f -
             inv2 * (0.0027777777777777778f - inv2 * 0.00079365079365079365f));
  const float stirling = (x - 0.5f) * lg - x + (0.91893853320467274f + series);
  return big ? stirling : small;
}

// FP32 log-gamma: vm_lgamma_pos for x > 0; for x <= 0 the reflection
// lgamma(x) = ln(pi / |sin(pi x)|) - lgamma(1 - x).  The hardware sine takes
// revolutions, sin(2 pi t), on a limited range, so t = x/2 is first reduced
// to [-1/2, 1/2] (exact: t - rint(t)); at the poles (integers x <= 0) the
// reduced argument is a whole half-revolution and the result +inf or very
// large.
__device__ __forceinline__ float vm_lgamma(float xin) {
  const bool refl = xin <= 0.0f;
  float r = vm_lgamma_pos(refl ? 1.0f - xin : xin);
  if (refl) {
    const float pi = 3.14159265358979f;
    float t = xin * 0.5f;
    t = t - rintf(t);
    const float s = __builtin_amdgcn_sinf(t);
    r = vm_log(fabsf(pi * vm_rcp(s))) - r;
  }
  return r;
}

// log(1 + exp(x)) without overflow.
__device__ __forceinline__ float vm_softplus(float x) {
  return fmaxf(x, 0.0f) + vm_log(1.0f + vm_exp(-fabsf(x)));
}

// log(exp(a) + exp(b)) that neither overflows nor underflows: the larger
// argument plus log(1 + exp(-|a - b|)), one exp and one log.  Equal
// arguments (equal infinities included) give a + ln 2, so two -inf terms
// stay -inf; a NaN argument propagates through the difference.
__device__ __forceinline__ float vm_lse(float a, float b) {
  const float m = fmaxf(a, b);
  const float dist = a == b ? 0.0f : fabsf(a - b);
  return m + vm_log(1.0f + vm_exp(-dist));
}

template <int OPT>
__device__ __forceinline__ void vm_load_col(const float* __restrict__ col,
                                            int lane, float (&v)[OPT]) {
  const float4* p = reinterpret_cast<const float4*>(col) + lane;
#pragma unroll
  for (int g = 0; g < OPT / 4; ++g) {
    float4 x = p[g * 64];
    v[4 * g + 0] = x.x;
    v[4 * g + 1] = x.y;
    v[4 * g + 2] = x.z;
    v[4 * g + 3] = x.w;
  }
}

template <int I, int NS, int OPT>
__device__ __forceinline__ void vm_spill(uint32_t slot, float (&S)[NS][OPT],
                                         const float (&T)[OPT]) {
  if constexpr (I < NS) {
    if (slot == I) {
#pragma unroll
      for (int k = 0; k < OPT; ++k) S[I][k] = T[k];
    } else {
      vm_spill<I + 1, NS, OPT>(slot, S, T);
    }
  }
}

template <int I, int NS, int OPT, typename F>
__device__ __forceinline__ void vm_bin_s(uint32_t slot, const float (&S)[NS][OPT],
                                         float (&T)[OPT], F f) {
  if constexpr (I < NS) {
    if (slot == I) {
#pragma unroll
      for (int k = 0; k < OPT; ++k) T[k] = f(S[I][k], T[k]);
    } else {
      vm_bin_s<I + 1, NS, OPT>(slot, S, T, f);
    }
  }
}

template <int I, int NS, int OPT>
__device__ __forceinline__ void vm_unspill(uint32_t slot,
                                           const float (&S)[NS][OPT],
                                           float (&T)[OPT]) {
  if constexpr (I < NS) {
    if (slot == I) {
#pragma unroll
      for (int k = 0; k < OPT; ++k) T[k] = S[I][k];
    } else {
      vm_unspill<I + 1, NS, OPT>(slot, S, T);
    }
  }
}

// Evaluate the program on one chunk; `tile` points at the chunk's first
// observation of column 0 and `ld` is the column stride of the tile (floats).
// The program lives in the constant address space: its address and the pc
// are wave-uniform, so every fetch is a scalar load into SGPRs.
typedef const __attribute__((address_space(4))) VmIns4* vm_prog_t;

__device__ __forceinline__ VmIns4 vm_fetch(vm_prog_t prog, uint32_t pc) {
  VmIns4 r;
  r.op = prog[pc].op;
  r.slot = prog[pc].slot;
  r.a = prog[pc].a;
  r.b = prog[pc].b;
  return r;
}

// One VM instruction on one chunk of CW chains (chain j's uniform table is
// U + j * nu).  An instruction that reads a data column reads it from LDS
// once for all CW chains; each chain's arithmetic is the single-chain
// arithmetic.  The interpreter calls it with the fetched opcode; code
// generated per model calls it with constants, so the switch folds away and
// each program becomes straight-line code.
template <int NS, int OPT, int CW>
__device__ __forceinline__ void vm_step(uint32_t op, uint32_t slot, uint32_t a,
                                        uint32_t b, const float* __restrict__ U,
                                        uint32_t nu,
                                        const float* __restrict__ tile,
                                        uint32_t ld, int lane,
                                        float (&T)[CW][OPT],
                                        float (&S)[CW][NS > 0 ? NS : 1][OPT],
                                        float (&D)[OPT]) {
  switch (op) {
    case VM_LOAD_U:
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const float u = U[j * nu + a];
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = u;
      }
      break;
    case VM_LOAD_D:
      vm_load_col<OPT>(tile + a * ld, lane, T[0]);
#pragma unroll
      for (int j = 1; j < CW; ++j)
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = T[0][k];
      break;
    case VM_PUSH_U:
      if constexpr (NS > 0) {
#pragma unroll
        for (int j = 0; j < CW; ++j) {
          vm_spill<0, NS, OPT>(slot, S[j], T[j]);
          const float u = U[j * nu + a];
#pragma unroll
          for (int k = 0; k < OPT; ++k) T[j][k] = u;
        }
      }
      break;
    case VM_PUSH_D:
      if constexpr (NS > 0) {
#pragma unroll
        for (int j = 0; j < CW; ++j) vm_spill<0, NS, OPT>(slot, S[j], T[j]);
        vm_load_col<OPT>(tile + a * ld, lane, T[0]);
#pragma unroll
        for (int j = 1; j < CW; ++j)
#pragma unroll
          for (int k = 0; k < OPT; ++k) T[j][k] = T[0][k];
      }
      break;
#define GPUM_SOP(NAME, EXPR)                                           \
case NAME:                                                             \
  if constexpr (NS > 0) {                                              \
    _Pragma("unroll") for (int j = 0; j < CW; ++j)                     \
        vm_bin_s<0, NS, OPT>(slot, S[j], T[j],                         \
                             [](float s, float t) { return (EXPR); }); \
  }                                                                    \
  break;
    GPUM_SOP(VM_ADD_S, s + t)
    GPUM_SOP(VM_SUB_S, s - t)
    GPUM_SOP(VM_MUL_S, s * t)
    GPUM_SOP(VM_DIV_S, s * vm_rcp(t))
    GPUM_SOP(VM_POW_S, vm_pow(s, t))
    GPUM_SOP(VM_LSE_S, vm_lse(s, t))
#undef GPUM_SOP
#define GPUM_UOP(NAME, EXPR)                              \
case NAME:                                              \
  _Pragma("unroll") for (int j = 0; j < CW; ++j) {      \
    const float u = U[j * nu + a];                      \
    _Pragma("unroll") for (int k = 0; k < OPT; ++k) {   \
      const float t = T[j][k];                          \
      T[j][k] = (EXPR);                                 \
    }                                                   \
  }                                                     \
  break;
    GPUM_UOP(VM_ADD_U, t + u)
    GPUM_UOP(VM_SUB_U, t - u)
    GPUM_UOP(VM_MUL_U, t * u)
    GPUM_UOP(VM_POW_U, vm_pow(t, u))
    GPUM_UOP(VM_RSUB_U, u - t)
    GPUM_UOP(VM_RDIV_U, u * vm_rcp(t))
    GPUM_UOP(VM_RPOW_U, vm_pow(u, t))
#undef GPUM_UOP
#define GPUM_DOP(NAME, EXPR)                              \
case NAME:                                              \
  vm_load_col<OPT>(tile + a * ld, lane, D);             \
  _Pragma("unroll") for (int j = 0; j < CW; ++j) {      \
    _Pragma("unroll") for (int k = 0; k < OPT; ++k) {   \
      const float t = T[j][k];                          \
      const float d = D[k];                             \
      T[j][k] = (EXPR);                                 \
    }                                                   \
  }                                                     \
  break;
    GPUM_DOP(VM_ADD_D, t + d)
    GPUM_DOP(VM_SUB_D, t - d)
    GPUM_DOP(VM_MUL_D, t * d)
    GPUM_DOP(VM_DIV_D, t * vm_rcp(d))
    GPUM_DOP(VM_POW_D, vm_pow(t, d))
    GPUM_DOP(VM_RSUB_D, d - t)
    GPUM_DOP(VM_RDIV_D, d * vm_rcp(t))
    GPUM_DOP(VM_RPOW_D, vm_pow(d, t))
#undef GPUM_DOP
#define GPUM_UNOP(NAME, EXPR)                             \
case NAME:                                              \
  _Pragma("unroll") for (int j = 0; j < CW; ++j) {      \
    _Pragma("unroll") for (int k = 0; k < OPT; ++k) {   \
      const float t = T[j][k];                          \
      T[j][k] = (EXPR);                                 \
    }                                                   \
  }                                                     \
  break;
    GPUM_UNOP(VM_NEG, -t)
    GPUM_UNOP(VM_EXP, vm_exp(t))
    GPUM_UNOP(VM_LOG, vm_log(t))
    GPUM_UNOP(VM_SQRT, vm_sqrt(t))
    GPUM_UNOP(VM_LGAMMA, vm_lgamma(t))
    GPUM_UNOP(VM_SQR, t * t)
    GPUM_UNOP(VM_CUBE, t * t * t)
    GPUM_UNOP(VM_POW4, (t * t) * (t * t))
    GPUM_UNOP(VM_RCP, vm_rcp(t))
    GPUM_UNOP(VM_SOFTPLUS, vm_softplus(t))
#undef GPUM_UNOP
    case VM_SAVE:
      if constexpr (NS > 0) {
#pragma unroll
        for (int j = 0; j < CW; ++j) vm_spill<0, NS, OPT>(slot, S[j], T[j]);
      }
      break;
    case VM_LOAD_S:
      if constexpr (NS > 0) {
#pragma unroll
        for (int j = 0; j < CW; ++j) vm_unspill<0, NS, OPT>(a, S[j], T[j]);
      }
      break;
    case VM_PUSH_S:
      if constexpr (NS > 0) {
#pragma unroll
        for (int j = 0; j < CW; ++j) {
          vm_spill<0, NS, OPT>(slot, S[j], T[j]);
          vm_unspill<0, NS, OPT>(a, S[j], T[j]);
        }
      }
      break;
    case VM_FMA_UU:
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const float ua = U[j * nu + a], ub = U[j * nu + b];
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = fmaf(T[j][k], ua, ub);
      }
      break;
    case VM_FMA_UD:
      vm_load_col<OPT>(tile + b * ld, lane, D);
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const float ua = U[j * nu + a];
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = fmaf(ua, D[k], T[j][k]);
      }
      break;
    case VM_FMA_DU:
      vm_load_col<OPT>(tile + a * ld, lane, D);
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const float ub = U[j * nu + b];
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = fmaf(T[j][k], D[k], ub);
      }
      break;
    default:
      break;
  }
}

// SUM_AFFINE: s[j] += U_j[a] * column b over the chunk, one fused operation
// per observation, in the order vm_load_col lays the observations out.
template <int OPT, int CW>
__device__ __forceinline__ void vm_acc_fma(const float* __restrict__ U,
                                           uint32_t nu,
                                           const float* __restrict__ tile,
                                           uint32_t ld, int lane, uint32_t a,
                                           uint32_t b, float (&s)[CW]) {
  float D[OPT];
  vm_load_col<OPT>(tile + b * ld, lane, D);
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    const float ua = U[j * nu + a];
#pragma unroll
    for (int k = 0; k < OPT; ++k) s[j] = fmaf(ua, D[k], s[j]);
  }
}

// vm_acc_fma with the chains' uniforms already in registers.
template <int OPT, int CW>
__device__ __forceinline__ void vm_acc_fma_u(const float (&ua)[CW],
                                             const float* __restrict__ tile,
                                             uint32_t ld, int lane, uint32_t b,
                                             float (&s)[CW]) {
  float D[OPT];
  vm_load_col<OPT>(tile + b * ld, lane, D);
#pragma unroll
  for (int j = 0; j < CW; ++j) {
#pragma unroll
    for (int k = 0; k < OPT; ++k) s[j] = fmaf(ua[j], D[k], s[j]);
  }
}

// A chunk of an affine program added to the sums s[j]: its FMA_UD
// instructions (pc >= 1; instruction 0 loads the zero the per-observation
// form starts from).  The caller starts s.
template <int OPT, int CW>
__device__ __forceinline__ void vm_run_acc(vm_prog_t prog, uint32_t n_instr,
                                           const float* __restrict__ U,
                                           uint32_t nu,
                                           const float* __restrict__ tile,
                                           uint32_t ld, int lane,
                                           float (&s)[CW]) {
  // Software-pipelined: the next product's column and uniforms are read
  // from LDS before this product's FMAs, and the instruction after it one
  // step earlier still (the fetch of index n_instr is the one the plain
  // loop also makes).  Each s[j] takes the same FMAs in the same order.
  if (n_instr < 2) return;
  VmIns4 cur = vm_fetch(prog, 1);
  VmIns4 nxt = vm_fetch(prog, 2);
  float D[OPT], ua[CW];
  vm_load_col<OPT>(tile + cur.b * ld, lane, D);
#pragma unroll
  for (int j = 0; j < CW; ++j) ua[j] = U[j * nu + cur.a];
  for (uint32_t pc = 1; pc + 1 < n_instr; ++pc) {
    const VmIns4 after = vm_fetch(prog, pc + 2);
    float Dn[OPT], uan[CW];
    vm_load_col<OPT>(tile + nxt.b * ld, lane, Dn);
#pragma unroll
    for (int j = 0; j < CW; ++j) uan[j] = U[j * nu + nxt.a];
#pragma unroll
    for (int j = 0; j < CW; ++j) {
#pragma unroll
      for (int k = 0; k < OPT; ++k) s[j] = fmaf(ua[j], D[k], s[j]);
    }
#pragma unroll
    for (int k = 0; k < OPT; ++k) D[k] = Dn[k];
#pragma unroll
    for (int j = 0; j < CW; ++j) ua[j] = uan[j];
    nxt = after;
  }
#pragma unroll
  for (int j = 0; j < CW; ++j) {
#pragma unroll
    for (int k = 0; k < OPT; ++k) s[j] = fmaf(ua[j], D[k], s[j]);
  }
}

// One entry of the sum tail (operands U[a], U[b]): the data pass's sum
// becomes alpha * s + n * beta.  One rounding per statement.
__device__ __forceinline__ void vm_tail_step(uint32_t op, double ua, double ub,
                                             double& alpha, double& beta) {
  if (op == VM_ADD_U) {
    beta = beta + ua;
  } else if (op == VM_SUB_U) {
    beta = beta - ua;
  } else {  // MUL_U; FMA_UU: T * U[a] + U[b]
    alpha = alpha * ua;
    beta = beta * ua;
    if (op == VM_FMA_UU) beta = beta + ub;
  }
}

// alpha * s + n * beta (n * beta only where the caller's sum covers the
// constant's share: once per data pass).
__device__ __forceinline__ double vm_tail_apply(double alpha, double beta,
                                                double s, uint32_t n,
                                                bool with_const) {
  const double nb = with_const ? (double)n * beta : 0.0;
  return __builtin_fma(alpha, s, nb);
}

template <int NS, int OPT, int CW>
__device__ __forceinline__ void vm_run(vm_prog_t prog, uint32_t n_instr,
                                       const float* __restrict__ U,
                                       uint32_t nu,
                                       const float* __restrict__ tile,
                                       uint32_t ld, int lane,
                                       float (&T)[CW][OPT]) {
  // NS == 0: a program of depth 1 (no PUSH, no binary op on the stack);
  // those cases compile to nothing and the spill registers do not exist.
  float S[CW][NS > 0 ? NS : 1][OPT];
  float D[OPT];
  // The first instruction of every program is a LOAD (the host lowering
  // starts from an empty stack), so T needs no initialisation.
  // Software-pipelined fetch: the next instruction's scalar load is issued
  // before the current one executes (the program carries one padding
  // instruction past its end).
  VmIns4 nxt = vm_fetch(prog, 0);
  for (uint32_t pc = 0; pc < n_instr; ++pc) {
    const uint32_t op = nxt.op, slot = nxt.slot, a = nxt.a, b = nxt.b;
    nxt = vm_fetch(prog, pc + 1);
    vm_step<NS, OPT, CW>(op, slot, a, b, U, nu, tile, ld, lane, T, S, D);
  }
}

}  // namespace gpum
