// lower.h -- host lowering of the reference bytecode into the HIP VM.
#pragma once
#include <stdint.h>

#include <string>
#include <vector>

#include "vm.h"

namespace gpum {

struct LowerOptions {
  // Evaluate data-only subtrees (e.g. log(y)) once on the host in FP64 and
  // stream them as extra FP32 columns instead of recomputing them per
  // iteration.
  bool precompute_data = true;
  // Fuse multiply-add patterns and reciprocal/negation of hoisted values,
  // and fold the program's tail into the summation (SumMode in vm.h).
  bool fuse = true;
};

struct PrecomputedColumn {
  // Postfix program in reference opcodes over the data row (PUSH_DATA args
  // index the original columns; PUSH_CONST args index `consts`).
  std::vector<uint32_t> code;
};

struct Lowered {
  std::vector<VmInstr> vm;  // per-observation register program
  int vm_depth = 0;         // peak values live (T + spill slots)
  // How the kernel sums the program over a chunk (SumMode), and the
  // ADD_U / SUB_U / MUL_U operations applied to a data pass's sum.
  uint32_t sum_mode = SUM_VALUE;
  std::vector<VmInstr> sum_tail;
  // FP64 program run once per proposal: reference opcodes over the
  // parameters plus OP_STORE_U* ops writing the uniform table.
  std::vector<uint32_t> hoist;
  std::vector<double> consts;  // constant pool of `hoist` and precomputed columns
  int n_uniform = 0;
  // Uniform entries that are constants: (slot, value) pairs written once.
  std::vector<std::pair<int, float>> const_uniforms;
  std::vector<PrecomputedColumn> extra_cols;
  int n_cols_in = 0;
  // VALU weights per observation evaluation (nominal slot counts):
  // w_exec counts the operations the kernel issues per observation besides
  // the summation's add (a summation fused with an operation counts as
  // that operation).
  double w_exec = 0.0;      // of the VM program the kernel executes
  double w_bytecode = 0.0;  // of the program as compiled by the DSL
  int ref_depth = 0;        // stack depth of the reference program
  bool empty = false;       // empty program: the data term is 0
};

// Validate a reference program.  `allow_data` is false for priors.
// Returns an empty string on success, else a message.
std::string validate_program(const uint32_t* code, size_t n_words,
                             size_t n_consts, uint32_t n_params,
                             uint32_t n_cols, bool allow_data, int* depth_out);

// Lower the log-likelihood program.  Returns "" on success.
std::string lower_program(const uint32_t* code, size_t n_words,
                          const double* consts, size_t n_consts,
                          uint32_t n_params, uint32_t n_cols,
                          const LowerOptions& opt, Lowered* out);

// FP64 evaluation of a reference program on one row (host helper used for
// precomputed columns; same semantics as the reference's f64 interpreter).
double eval_ref64(const uint32_t* code, size_t n_words, const double* consts,
                  const double* params, const double* row);

// A parameter program (hoist or prior, reference opcodes) with every
// `PUSH_CONST c; POW` whose constant c is 2 replaced by OP_SQUARE: x * x is
// the correctly rounded value of pow(x, 2), without the general routine.
std::vector<uint32_t> square_pow2(const uint32_t* code, size_t n_words,
                                  const double* consts);

// Human-readable listing of a lowered program (debugging aid).
std::string disassemble(const Lowered& L);

}  // namespace gpum
