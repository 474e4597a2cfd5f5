// eval64.cpp -- FP64 evaluation of a log-likelihood program at many
// parameter points on the device: the per-observation matrix behind WAIC /
// PSIS-LOO and the gradient behind the Cramer-Rao diagnostic
// (cpu_native.rs:306-329, 371-391; their FFI lib.rs:158-207).
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "eval64.h"
#include "jit.h"
#include "runtime.h"

namespace gpum {
namespace {

using rt::DevBuf;
using rt::g_dev;
using rt::set_err;
using rt::set_hip_err;

thread_local double g_last_eval_ms = 0.0;
thread_local int g_last_eval_compiled = 0;
// Kernels compiled per program when the process can (else, and when this
// is off, the interpreter).
bool g_eval_jit = true;

// Device buffers bounded per launch: a pointwise chunk's output and a
// gradient chunk's partial sums.
constexpr size_t kPointwiseChunkBytes = size_t(1) << 30;
constexpr size_t kGradPartialBytes = size_t(64) << 20;

bool is_leaf(uint32_t op) { return op <= OP_PUSH_DATA; }
bool is_unary_op(uint32_t op) {
  return op == OP_NEG || op == OP_EXP || op == OP_LOG || op == OP_SQRT ||
         op == OP_LGAMMA;
}

}  // namespace

hipFunction_t eval64_compiled_kernel(const uint32_t* code, size_t code_len,
                                     const double* consts, int kind) {
  if (!g_eval_jit || code_len == 0) return nullptr;
  std::string err;
  return jit_kernel(jit_eval64_source(code, code_len, consts), (JitKind)kind,
                    &err);
}

// The program with each instruction's operand positions: a stack of the
// instructions whose values are live, as the reference's forward pass keeps
// (cpu_native.rs:156-211).  Then the values the backward pass reads
// (cpu_native.rs:218-277) get slots: the operands of MUL, the divisor of
// DIV, the operands of LOG, LGAMMA and POW, and the results of DIV, EXP,
// SQRT and POW, unless they are leaves (re-read instead).
std::vector<TapeOp> make_tape(const uint32_t* code, size_t n_ops,
                              uint32_t* n_slots) {
  std::vector<TapeOp> tape(n_ops);
  std::vector<uint32_t> st;
  for (size_t pc = 0; pc < n_ops; ++pc) {
    TapeOp o{code[2 * pc], code[2 * pc + 1], 0u, 0u, kNoSlot};
    if (!is_leaf(o.op)) {
      if (is_unary_op(o.op)) {
        o.l = st.back();
        st.pop_back();
      } else {
        o.r = st.back();
        st.pop_back();
        o.l = st.back();
        st.pop_back();
      }
    }
    st.push_back((uint32_t)pc);
    tape[pc] = o;
  }
  std::vector<char> keep(n_ops, 0);
  for (size_t pc = 0; pc < n_ops; ++pc) {
    const TapeOp& o = tape[pc];
    switch (o.op) {
      case OP_MUL: keep[o.l] = keep[o.r] = 1; break;
      case OP_DIV: keep[o.r] = 1; keep[pc] = 1; break;
      case OP_LOG:
      case OP_LGAMMA: keep[o.l] = 1; break;
      case OP_EXP:
      case OP_SQRT: keep[pc] = 1; break;
      case OP_POW: keep[o.l] = keep[o.r] = 1; keep[pc] = 1; break;
      default: break;
    }
  }
  uint32_t n = 0;
  for (size_t pc = 0; pc < n_ops; ++pc)
    if (keep[pc] && !is_leaf(tape[pc].op)) tape[pc].slot = n++;
  if (n_slots) *n_slots = n;
  return tape;
}

namespace {

// Everything one call uploads.
struct EvalSetup {
  DevBuf<TapeOp> prog;
  DevBuf<double> consts, data, points;
  Eval64Args a{};
};

// The program's checks (validate_program) plus its peak stack depth.
bool check_program(const uint32_t* code, size_t code_len, size_t n_consts,
                   uint32_t n_params, uint32_t n_cols, int* depth) {
  if (code_len > 0 && !code) {
    set_err("code must not be NULL");
    return false;
  }
  std::string e = validate_program(code, code_len, n_consts, n_params, n_cols,
                                   true, depth);
  if (!e.empty()) {
    set_err("%s", e.c_str());
    return false;
  }
  return true;
}

bool setup(const uint32_t* code, size_t code_len, const double* consts,
           size_t n_consts, uint32_t n_params, const double* data,
           uint32_t n_cols, uint32_t n_obs, const double* points,
           uint32_t n_points, int depth, EvalSetup* S) {
  if (n_consts > 0 && !consts) {
    set_err("consts must not be NULL");
    return false;
  }
  if ((size_t)n_obs * n_cols > 0 && !data) {
    set_err("data must not be NULL");
    return false;
  }
  const size_t n_ops = code_len / 2;
  uint32_t n_slots = 0;
  std::vector<TapeOp> tape = make_tape(code, n_ops, &n_slots);
  // Row-major (n_obs x n_cols) in, column-major on the device.
  std::vector<double> cols((size_t)n_obs * n_cols);
  for (uint32_t i = 0; i < n_obs; ++i)
    for (uint32_t c = 0; c < n_cols; ++c)
      cols[(size_t)c * n_obs + i] = data[(size_t)i * n_cols + c];
  if (!S->prog.upload(tape.data(), tape.size(), "uploading the program") ||
      !S->consts.upload(consts, n_consts, "uploading constants") ||
      !S->data.upload(cols.data(), cols.size(), "uploading data") ||
      !S->points.upload(points, (size_t)n_points * n_params,
                        "uploading points"))
    return false;
  Eval64Args& a = S->a;
  a.prog = S->prog.p;
  a.n_ops = (uint32_t)n_ops;
  a.n_params = n_params;
  a.consts = S->consts.p;
  a.data = S->data.p;
  a.n_obs = n_obs;
  a.n_points = n_points;
  a.points = S->points.p;
  a.depth = (uint32_t)depth;
  a.n_slots = n_slots;
  return true;
}

uint32_t grid_for(uint64_t items) {
  const uint64_t cap = (uint64_t)g_dev.n_cu * 8u;
  return (uint32_t)std::max<uint64_t>(1, std::min(items, cap));
}

// Kernel time between two events already recorded on the stream.
double elapsed_ms(hipEvent_t e0, hipEvent_t e1) {
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  return (double)ms;
}

hipFunction_t compiled(const uint32_t* code, size_t code_len,
                       const double* consts, JitKind kind) {
  return eval64_compiled_kernel(code, code_len, consts, kind);
}

hipError_t launch_compiled(hipFunction_t f, const Eval64Args& a, uint32_t grid,
                           hipStream_t st) {
  Eval64Args arg = a;
  void* kargs[] = {(void*)&arg};
  return rt::launch_module(f, dim3(grid), dim3(256), 0, st, kargs);
}

struct Events {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Events() {
    hipEventCreate(&e0);
    hipEventCreate(&e1);
  }
  ~Events() {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
  }
};

}  // namespace
}  // namespace gpum

using namespace gpum;

extern "C" {

int gpum_hip_loglik_pointwise(const uint32_t* code, size_t code_len,
                              const double* consts, size_t n_consts,
                              uint32_t n_params, const double* data,
                              uint32_t n_cols, uint32_t n_obs,
                              const double* points, uint32_t n_points,
                              double* out) {
  std::lock_guard<std::mutex> lk(rt::g_mu);
  g_last_eval_ms = 0.0;
  g_last_eval_compiled = 0;
  if (!rt::device_init()) return 1;
  hipSetDevice(g_dev.id);
  int depth = 0;
  if (!check_program(code, code_len, n_consts, n_params, n_cols, &depth))
    return 1;
  // As the reference: the point count is points.len() / n_params.
  if (n_params == 0) n_points = 0;
  if (n_points == 0 || n_obs == 0) return 0;
  if (!points || !out) {
    set_err("points and out must not be NULL");
    return 1;
  }
  if (code_len == 0) {
    // An empty program evaluates to 0 (cpu_native.rs:82-84).
    memset(out, 0, (size_t)n_points * n_obs * sizeof(double));
    return 0;
  }
  EvalSetup S;
  if (!setup(code, code_len, consts, n_consts, n_params, data, n_cols, n_obs,
             points, n_points, depth, &S))
    return 1;
  const uint32_t chunk = (uint32_t)std::max<size_t>(
      1, std::min<size_t>(n_points,
                          kPointwiseChunkBytes / ((size_t)n_obs * 8)));
  DevBuf<double> d_out;
  if (!d_out.alloc((size_t)chunk * n_obs, "allocating the output")) return 1;
  hipStream_t st = g_dev.stream;
  const hipFunction_t jf = compiled(code, code_len, consts, JIT_POINTWISE64);
  g_last_eval_compiled = jf ? 1 : 0;
  Events ev;
  double ms = 0.0;
  for (uint32_t p0 = 0; p0 < n_points; p0 += chunk) {
    const uint32_t np = std::min(chunk, n_points - p0);
    Eval64Args a = S.a;
    a.points = S.points.p + (size_t)p0 * n_params;
    a.n_points = np;
    a.out = d_out.p;
    // Work items: (point, slice) on the interpreter, (group of kPointGroup
    // points, slice) on the compiled kernel.
    const uint64_t blocks =
        jf ? (uint64_t)((n_obs + kPointwiseSliceObs - 1) / kPointwiseSliceObs) *
                 ((np + kPointGroup - 1) / kPointGroup)
           : (uint64_t)((n_obs + kGradSliceObs - 1) / kGradSliceObs) * np;
    hipError_t e = hipEventRecord(ev.e0, st);
    if (e == hipSuccess)
      e = jf ? launch_compiled(jf, a, grid_for(blocks), st)
             : launch_pointwise64(a, grid_for(blocks), st);
    if (e == hipSuccess) e = hipEventRecord(ev.e1, st);
    if (e == hipSuccess)
      e = hipMemcpyAsync(out + (size_t)p0 * n_obs, d_out.p,
                         (size_t)np * n_obs * sizeof(double),
                         hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_hip_err("pointwise log-likelihood", e);
      return 1;
    }
    ms += elapsed_ms(ev.e0, ev.e1);
  }
  g_last_eval_ms = ms;
  return 0;
}

int gpum_hip_grad_batch(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        uint32_t n_params, const double* data, uint32_t n_cols,
                        uint32_t n_obs, const double* points, uint32_t n_points,
                        double* out) {
  std::lock_guard<std::mutex> lk(rt::g_mu);
  g_last_eval_ms = 0.0;
  g_last_eval_compiled = 0;
  if (!rt::device_init()) return 1;
  hipSetDevice(g_dev.id);
  int depth = 0;
  if (!check_program(code, code_len, n_consts, n_params, n_cols, &depth))
    return 1;
  if (n_params == 0) n_points = 0;
  if (n_points == 0) return 0;
  if (!points || !out) {
    set_err("points and out must not be NULL");
    return 1;
  }
  if (n_params > (uint32_t)kMaxParams) {
    set_err("the gradient supports at most %d parameters", kMaxParams);
    return 1;
  }
  if (code_len / 2 > kMaxTapeOps) {
    set_err("the gradient supports programs of at most %u instructions",
            kMaxTapeOps);
    return 1;
  }
  if (code_len == 0 || n_obs == 0) {
    // No term depends on the parameters: the gradient is zero.
    memset(out, 0, (size_t)n_points * n_params * sizeof(double));
    return 0;
  }
  EvalSetup S;
  if (!setup(code, code_len, consts, n_consts, n_params, data, n_cols, n_obs,
             points, n_points, depth, &S))
    return 1;
  const uint32_t ns =
      (uint32_t)(((uint64_t)n_obs + kGradSliceObs - 1) / kGradSliceObs);
  const uint32_t chunk = (uint32_t)std::max<size_t>(
      1, std::min<size_t>(n_points,
                          kGradPartialBytes / ((size_t)ns * (n_params + 1) * 8)));
  DevBuf<double> d_part, d_out;
  if (!d_part.alloc((size_t)chunk * ns * (n_params + 1),
                    "allocating partial sums") ||
      !d_out.alloc((size_t)n_points * n_params, "allocating the output"))
    return 1;
  hipStream_t st = g_dev.stream;
  const hipFunction_t jf = compiled(code, code_len, consts, JIT_GRAD64);
  g_last_eval_compiled = jf ? 1 : 0;
  Events ev;
  hipError_t e = hipEventRecord(ev.e0, st);
  for (uint32_t p0 = 0; e == hipSuccess && p0 < n_points; p0 += chunk) {
    const uint32_t np = std::min(chunk, n_points - p0);
    Eval64Args a = S.a;
    a.points = S.points.p + (size_t)p0 * n_params;
    a.n_points = np;
    a.out = d_part.p;
    a.n_slices = ns;
    e = jf ? launch_compiled(jf, a, grid_for(grad64_items(ns, np)), st)
           : launch_grad64(a, grid_for(grad64_items(ns, np)), st);
    if (e == hipSuccess)
      e = launch_grad64_finish(d_part.p, np, n_params, ns,
                               d_out.p + (size_t)p0 * n_params, nullptr, st);
  }
  if (e == hipSuccess) e = hipEventRecord(ev.e1, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out, d_out.p, (size_t)n_points * n_params * sizeof(double),
                       hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_hip_err("gradient batch", e);
    return 1;
  }
  g_last_eval_ms = elapsed_ms(ev.e0, ev.e1);
  return 0;
}

double gpum_hip_last_eval_ms(void) { return g_last_eval_ms; }

int gpum_hip_last_eval_compiled(void) { return g_last_eval_compiled; }

void gpum_hip_set_eval_jit(int32_t enable) {
  std::lock_guard<std::mutex> lk(rt::g_mu);
  g_eval_jit = enable != 0;
}

}  // extern "C"
