// eval64_device.h -- FP64 evaluation of one reference program at one
// observation on the device: the stack machine (cpu_native.rs:80-107) and
// value plus reverse-mode gradient through a per-lane tape
// (cpu_native.rs:148-279).  Shared by the batch-evaluation kernels and the
// Langevin sampler.  Every operation is rounded as the reference's: the
// function bodies turn off fused multiply-add.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "eval64.h"
#include "math64.h"
#include "vm.h"

namespace gpum {

__device__ __forceinline__ double data_at(const Eval64Args& a, uint32_t col,
                                          uint32_t i) {
  return a.data[(size_t)col * a.n_obs + i];
}

// a^b as the reference's powf: a * a for b = 2 and a for b = 1 (the
// correctly rounded power, special values included), the library routine
// for every other exponent.  The exponent of a power and of its derivative
// (b - 1) is most often the constant 2.  The empty asm hides a constant
// exponent from the optimiser, so that code generated per program calls
// the routine the interpreter calls rather than a rewrite (pow(x, -1) as
// 1 / x, say).
__device__ __forceinline__ double pow_ref64(double a, double b) {
  if (b == 2.0) return a * a;
  if (b == 1.0) return a;
  __asm__ volatile("" : "+v"(b));
  return pow(a, b);
}

// The stack machine, cpu_native.rs:80-107.  The stack pointer follows the
// program only, so it is the same in every lane.
template <int MAXD>
__device__ double eval_obs64(const Eval64Args& a, const double* P,
                             uint32_t i) {
#pragma clang fp contract(off)
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
      default: st[sp - 2] = pow_ref64(st[sp - 2], st[sp - 1]); --sp; break;
    }
  }
  return st[0];
}

// A leaf's value: the constant, the parameter or the observation's datum.
__device__ __forceinline__ double leaf64(const Eval64Args& a, const TapeOp& o,
                                         const double* P, uint32_t i) {
  if (o.op == OP_PUSH_CONST) return a.consts[o.arg];
  if (o.op == OP_PUSH_PARAM) return P[o.arg];
  return data_at(a, o.arg, i);
}

// One observation's value and gradient, cpu_native.rs:148-279.  The forward
// pass is the stack machine; the values the backward pass reads are kept in
// their slots (leaves are re-read instead).  Each value is the operand of
// exactly one later instruction, so its adjoint receives exactly one
// contribution: walking the program backwards, the pending adjoints form a
// stack (the operands' adjoints are pushed, the right operand's last, and
// the right operand is the instruction just before).  A zero adjoint
// propagates zeros, as the reference skips it.  d/d param k accumulates
// into g[k]; returns the value.  An empty program is not evaluated here.
// MAXD bounds the stack depth, MAXS the kept values.
template <int MAXD, int MAXS>
__device__ double grad_obs64(const Eval64Args& a, const double* P, uint32_t i,
                             double (&g)[kMaxParams]) {
#pragma clang fp contract(off)
  double st[MAXD];
  double kept[MAXS];
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
      default: st[sp - 2] = pow_ref64(st[sp - 2], st[sp - 1]); --sp; break;
    }
    if (o.slot != kNoSlot) kept[o.slot] = st[sp - 1];
  }
  const double value = st[0];
  // An operand's value in the backward pass.
  auto val = [&](uint32_t k) -> double {
    const TapeOp q = a.prog[k];
    return q.op <= OP_PUSH_DATA ? leaf64(a, q, P, i) : kept[q.slot];
  };
  double adj[MAXD];
  sp = 0;
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
      case OP_MUL: dl = ga * val(o.r); dr = ga * val(o.l); break;
      case OP_DIV: {
        const double b = val(o.r);
        dl = ga / b;
        dr = -(ga * kept[o.slot] / b);
        break;
      }
      case OP_NEG: dl = -ga; break;
      case OP_EXP: dl = ga * kept[o.slot]; break;
      case OP_LOG: dl = ga / val(o.l); break;
      case OP_SQRT: dl = ga * 0.5 / kept[o.slot]; break;
      case OP_LGAMMA: dl = ga * digamma64(val(o.l)); break;
      default: {
        const double base = val(o.l), ex = val(o.r);
        dl = ga * ex * pow_ref64(base, ex - 1.0);
        dr = base > 0.0 ? ga * kept[o.slot] * log(base) : 0.0;
        break;
      }
    }
    if (o.op >= OP_ADD) {
      adj[sp++] = live ? dl : 0.0;
      if (o.op <= OP_DIV || o.op == OP_POW) adj[sp++] = live ? dr : 0.0;
    }
  }
  return value;
}

}  // namespace gpum
