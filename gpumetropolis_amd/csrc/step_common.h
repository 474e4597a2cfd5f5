// step_common.h -- FP64 per-chain step code shared by the sampler kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "math64.h"
#include "philox.h"
#include "vm.h"

namespace gpum {

// Hides a compile-time constant from the optimiser: code generated per
// model passes its constant exponents through it, so pow(x, 2.0) there
// calls the same library routine the interpreter calls, not a rewrite.
__device__ __forceinline__ double opaque64(double x) {
  __asm__ volatile("" : "+v"(x));
  return x;
}

// FP64 stack evaluation of a reference program over the parameters (no
// data), with the hoist extensions writing the uniform table.  Code
// generated per model (jit.cpp) performs the same operations in the same
// order.
static __device__ __noinline__ double eval_params64(const uint32_t* __restrict__ code,
                                uint32_t n_words,
                                const double* __restrict__ consts,
                                const double* params, float* U) {
  if (n_words < 2) return 0.0;
  double st[kMaxStack + 1];
  int sp = 0;
  for (uint32_t pc = 0; pc < n_words / 2; ++pc) {
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    switch (op) {
      case OP_PUSH_CONST: st[sp++] = consts[arg]; break;
      case OP_PUSH_PARAM: st[sp++] = params[arg]; break;
      case OP_ADD: st[sp - 2] += st[sp - 1]; --sp; break;
      case OP_SUB: st[sp - 2] -= st[sp - 1]; --sp; break;
      case OP_MUL: st[sp - 2] *= st[sp - 1]; --sp; break;
      case OP_DIV: st[sp - 2] /= st[sp - 1]; --sp; break;
      case OP_NEG: st[sp - 1] = -st[sp - 1]; break;
      case OP_EXP: st[sp - 1] = exp(st[sp - 1]); break;
      case OP_LOG: st[sp - 1] = log(st[sp - 1]); break;
      case OP_SQRT: st[sp - 1] = sqrt(st[sp - 1]); break;
      case OP_LGAMMA: st[sp - 1] = lgamma64(st[sp - 1]); break;
      case OP_POW: st[sp - 2] = pow(st[sp - 2], st[sp - 1]); --sp; break;
      case OP_STORE_U: U[arg] = (float)st[--sp]; break;
      case OP_STORE_U_RCP: U[arg] = (float)(1.0 / st[--sp]); break;
      case OP_STORE_U_NEG: U[arg] = (float)(-st[--sp]); break;
      case OP_SQUARE: st[sp - 1] = st[sp - 1] * st[sp - 1]; break;
      default: break;
    }
  }
  return sp > 0 ? st[0] : 0.0;
}

// The uniforms of one chain-iteration, materialised four at a time.
struct StepUniforms {
  uint32_t seed, stream, t, chain;
  U4 cache;
  int cached = -1;
  __device__ double get(uint32_t slot) {
    int blk = (int)(slot >> 2);
    if (blk != cached) {
      cache = philox4x32_10(U4{t, chain, (uint32_t)blk, TAG_STEP}, seed, stream);
      cached = blk;
    }
    return u32_to_unit(u4_word(cache, slot & 3u));
  }
};

// Target acceptance by dimension (R/adaptive.R:69-72).
static __device__ __forceinline__ double target_accept(int d) {
  const double low[8] = {0.44, 0.35, 0.32, 0.30, 0.28, 0.27, 0.26, 0.25};
  return (d >= 1 && d <= 8) ? low[d - 1] : 0.234;
}

}  // namespace gpum
