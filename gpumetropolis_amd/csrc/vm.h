// vm.h -- the per-observation virtual machine the HIP interpreter executes.
//
// The reference ships a 13-opcode postfix stack machine (R/dsl.R:13-17,
// interpreted per observation by chain_kernel.rs:61-120 and
// cpu_native.rs:80-107).  The HIP backend accepts that bytecode unchanged at
// its boundary and lowers it on the host (lower.cpp) into this VM:
//
//   * parameter-only subtrees are hoisted out of the observation loop and
//     evaluated once per proposal in FP64 (the "hoist" program, written in
//     the reference's own opcodes plus the STORE ops below);
//   * constants and hoisted values become wave-uniform operands U[k]
//     (FP32, one table per chain, in LDS);
//   * what is left -- the part that depends on the data -- becomes a short
//     register program over a cached top-of-stack T and spill slots S[i].
//     The stack pointer of a postfix program is static per pc, so every
//     operand's slot is known on the host; only PUSH and the binary-on-stack
//     ops carry a slot, and the kernel resolves it with a small switch so
//     the register file stays statically indexed.
//
// Each lane evaluates the program on OPT observations at once, so one
// dispatch of an opcode (wave-uniform, fetched with a scalar load) drives
// OPT vector operations.
#pragma once
#include <stdint.h>

namespace gpum {

// Limits of the reference's kernel (chain_kernel.rs:18-19, R/dsl.R:20).
constexpr int kMaxStack = 32;
constexpr int kMaxParams = 16;
// Uniform operand table per chain (hoisted values + constants).
constexpr int kMaxUniform = 256;
// Maximum VM program length (instructions).
constexpr int kMaxVmInstr = 1024;
// Maximum data columns (original + precomputed).
constexpr int kMaxCols = 64;

enum VmOp : uint32_t {
  // T <- operand (stack was empty)
  VM_LOAD_U = 0,
  VM_LOAD_D,
  // S[slot] <- T ; T <- operand
  VM_PUSH_U,
  VM_PUSH_D,
  // T <- S[slot] op T   (pops one)
  VM_ADD_S,
  VM_SUB_S,
  VM_MUL_S,
  VM_DIV_S,
  VM_POW_S,
  // T <- T op U[a]
  VM_ADD_U,
  VM_SUB_U,
  VM_MUL_U,
  VM_POW_U,
  // T <- U[a] op T
  VM_RSUB_U,
  VM_RDIV_U,
  VM_RPOW_U,
  // T <- T op D[a]
  VM_ADD_D,
  VM_SUB_D,
  VM_MUL_D,
  VM_DIV_D,
  VM_POW_D,
  // T <- D[a] op T
  VM_RSUB_D,
  VM_RDIV_D,
  VM_RPOW_D,
  // unary on T
  VM_NEG,
  VM_EXP,
  VM_LOG,
  VM_SQRT,
  VM_LGAMMA,
  VM_SQR,
  VM_CUBE,
  VM_POW4,
  VM_RCP,
  // fused
  VM_FMA_UU,  // T <- T * U[a] + U[b]
  VM_FMA_UD,  // T <- T + U[a] * D[b]
  VM_FMA_DU,  // T <- T * D[a] + U[b]
  VM_SOFTPLUS,  // T <- log(1 + exp(T)), overflow-safe
  // common subexpressions kept in spill slots above the program's stack
  VM_SAVE,    // S[slot] <- T
  VM_LOAD_S,  // T <- S[a] (stack was empty)
  VM_PUSH_S,  // S[slot] <- T ; T <- S[a]
  // T <- log(exp(S[slot]) + exp(T)) without underflow (pops one)
  VM_LSE_S,
  VM_NUM_OPS
};

// How a chunk of observations is summed.  The kernels need the sum of the
// program's value over the observations, not the values, so with fusion
// the lowering moves work into that sum (lower.cpp, fold_sum):
//   SUM_VALUE   s += T
//   SUM_SQUARE  s = fma(T, T, s): the program's trailing SQR
//   SUM_AFFINE  the program is c + sum_i U[a_i] * D[b_i]: no T at all,
//               s = fma(U[a_i], D[b_i], s) for every FMA_UD of the
//               program (its first instruction loads a zero) and c joins
//               the tail below.
// Trailing ADD_U / SUB_U / MUL_U / FMA_UU leave the per-observation program
// and are applied once to the data pass's sum (the "sum tail"), in FP64:
//   sum(T + u) = sum(T) + n u,   sum(u T) = u sum(T).
enum SumMode : uint32_t { SUM_VALUE = 0, SUM_SQUARE = 1, SUM_AFFINE = 2 };

// One VM instruction: two u32 words.
//   w0 = op | slot << 8
//   w1 = a | b << 16
struct VmInstr {
  uint32_t w0;
  uint32_t w1;
};

// The decoded form the kernel fetches (one 16-byte scalar load).
struct alignas(16) VmIns4 {
  uint32_t op, slot, a, b;
};

inline uint32_t vm_op(VmInstr i) { return i.w0 & 0xFFu; }
inline uint32_t vm_slot(VmInstr i) { return i.w0 >> 8; }
inline uint32_t vm_a(VmInstr i) { return i.w1 & 0xFFFFu; }
inline uint32_t vm_b(VmInstr i) { return i.w1 >> 16; }

// Opcodes of the reference bytecode (R/dsl.R:13-17).
enum RefOp : uint32_t {
  OP_PUSH_CONST = 0,
  OP_PUSH_PARAM = 1,
  OP_PUSH_DATA = 2,
  OP_ADD = 3,
  OP_SUB = 4,
  OP_MUL = 5,
  OP_DIV = 6,
  OP_NEG = 7,
  OP_EXP = 8,
  OP_LOG = 9,
  OP_SQRT = 10,
  OP_POW = 11,
  OP_LGAMMA = 12,
  // Hoist-program extensions (never accepted at the boundary):
  OP_STORE_U = 100,      // pop -> U[arg] = (float)v
  OP_STORE_U_RCP = 101,  // pop -> U[arg] = (float)(1 / v)
  OP_STORE_U_NEG = 102,  // pop -> U[arg] = (float)(-v)
  // x^2 with a constant exponent 2, rewritten by square_pow2 (lower.cpp):
  OP_SQUARE = 103,       // top -> top * top (what a correctly rounded pow gives)
};

}  // namespace gpum
