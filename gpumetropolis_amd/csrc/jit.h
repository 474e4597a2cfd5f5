// jit.h -- per-model kernels compiled at run time (hiprtc).
//
// The reference compiles each model's log-likelihood to native code for its
// CPU path (src/rust/src/jit.rs:441-631, Cranelift).  This is the same idea
// on the GPU: the lowered VM program becomes a straight-line evaluator
// (the interpreter's own vm_step calls with constant operands, so the
// arithmetic is the interpreter's, operation for operation) instantiated
// into the chain, points and batched kernel bodies.  Modules are cached per
// process by source text, so models sharing a program share kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

namespace gpum {

struct Lowered;

enum JitKind { JIT_CHAIN_STREAM = 0, JIT_CHAIN_RESIDENT, JIT_POINTS_STREAM,
               JIT_POINTS_RESIDENT, JIT_BATCH, JIT_POINTWISE64, JIT_GRAD64,
               JIT_NUM_KINDS };
// Kinds compiled from jit_evaluator_source; JIT_POINTWISE64 and JIT_GRAD64
// from jit_eval64_source.
constexpr int kJitChainKinds = JIT_POINTWISE64;

// Chunks of 64*opt observations that one FP32 partial covers for a program
// of `ns` spill slots over rows of n_cols columns: two when the program has
// no spill slots and two chunks of rows fit one LDS tile buffer (kTileF4
// float4s), else one.  The tiling and the compiled evaluator both use it, so
// a model's observations are grouped alike on every path.
uint32_t unit_chunks_for(int ns, int opt, uint32_t n_cols);

// Source of the evaluator for a lowered program run with variant
// (ns, opt, cw): ns spill slots, opt observations per lane per chunk, cw
// chains per wave; with the model's FP64 parameter programs (the lowered
// hoist program and the prior, reference bytecode over the parameters).
std::string jit_evaluator_source(const Lowered& L, const uint32_t* prior,
                                 size_t n_prior, const double* prior_consts,
                                 int ns, int opt, int cw);

// Source of the FP64 evaluator of a reference program (valid,
// validate_program): its value and, by reverse mode, its gradient as
// straight-line code, every operation rounded as the interpreter's
// (eval64_device.h).  Instantiated into the bodies of eval64_kernels.h.
std::string jit_eval64_source(const uint32_t* code, size_t n_words,
                              const double* consts);

// The kernel of `kind` for an evaluator source (compiled on first use).
// On failure returns nullptr and fills *err.
hipFunction_t jit_kernel(const std::string& evaluator, JitKind kind,
                         std::string* err, double* compile_ms = nullptr);

// Compile the kernel of `kind` with hiprtc without loading it: needs no
// device, only the compiler.  On failure fills *err.
bool jit_compile_only(const std::string& evaluator, JitKind kind,
                      std::string* err, double* compile_ms = nullptr);

// Dynamic LDS a module kernel may use without an attribute call.
constexpr size_t kJitMaxDynamicLds = 65536;

}  // namespace gpum
