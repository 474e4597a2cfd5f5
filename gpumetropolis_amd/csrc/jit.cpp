// jit.cpp -- compile per-model kernels with hiprtc and cache them.
#include "jit.h"

#include <hip/hip_runtime.h>
#include <hip/hip_version.h>
#include <hip/hiprtc.h>

#include <string.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "chain_kernel.h"
#include "lower.h"
#include "vm.h"

// kJitHeaderNames / kJitHeaderSources / kJitHeaderCount: the device headers
// the kernels are built from, embedded at build time.
#include "jit_sources.inc"

namespace gpum {
namespace {

// hiprtc predefines the HIP runtime and the device math functions but ships
// no system headers; these stand in for the three the device headers name.
const char* kStubRuntime = "#pragma once\n";
const char* kStubMath = "#pragma once\n";
const char* kStubStdint =
    "#pragma once\n"
    "typedef __hip_internal::uint8_t uint8_t;\n"
    "typedef __hip_internal::uint16_t uint16_t;\n"
    "typedef __hip_internal::uint32_t uint32_t;\n"
    "typedef __hip_internal::uint64_t uint64_t;\n"
    "typedef __hip_internal::int8_t int8_t;\n"
    "typedef __hip_internal::int16_t int16_t;\n"
    "typedef __hip_internal::int32_t int32_t;\n"
    "typedef __hip_internal::int64_t int64_t;\n";

const char* kKernels[JIT_NUM_KINDS] = {
    "extern \"C\" __global__ void __launch_bounds__(1024)\n"
    "gpum_jit_kernel(const gpum::ChainArgs* A, gpum::ChainHot H) {\n"
    "  gpum::chain_body<gpum::JitEval, true>(A, H, blockIdx.x * H.cpb);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(1024)\n"
    "gpum_jit_kernel(const gpum::ChainArgs* A, gpum::ChainHot H) {\n"
    "  gpum::chain_body<gpum::JitEval, false>(A, H, blockIdx.x * H.cpb);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(1024)\n"
    "gpum_jit_kernel(const gpum::PointsArgs* P, gpum::ChainHot H) {\n"
    "  gpum::points_body<gpum::JitEval, true>(P, H);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(1024)\n"
    "gpum_jit_kernel(const gpum::PointsArgs* P, gpum::ChainHot H) {\n"
    "  gpum::points_body<gpum::JitEval, false>(P, H);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(1024)\n"
    "gpum_jit_kernel(const gpum::BatchQueue* q, const gpum::ChainArgs* a,\n"
    "                const gpum::ChainHot* h) {\n"
    "  gpum::batch_body<gpum::JitEval>(q, a, h);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(256)\n"
    "gpum_jit_kernel(gpum::Eval64Args a) {\n"
    "  gpum::pointwise64_grouped_body<gpum::JitEval64>(a);\n}\n",
    "extern \"C\" __global__ void __launch_bounds__(256, 4)\n"
    "gpum_jit_kernel(gpum::Eval64Args a) {\n"
    "  gpum::grad64_body<gpum::JitEval64>(a);\n}\n",
};

// The gradient kernel above runs four waves per SIMD (at most 128 registers
// a lane); a program whose kernel then keeps more than kGradSpillBytes a
// lane in scratch memory is compiled again for two waves per SIMD.
const char* kGrad64TwoWaves =
    "extern \"C\" __global__ void __launch_bounds__(256, 2)\n"
    "gpum_jit_kernel(gpum::Eval64Args a) {\n"
    "  gpum::grad64_body<gpum::JitEval64>(a);\n}\n";
constexpr int kGradSpillBytes = 64;

struct Entry {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

std::mutex g_jit_mu;
std::map<std::string, Entry> g_cache;  // by full source text

// hiprtc's code object for a kernel source (no device involved).
bool compile_code(const std::string& src, std::vector<char>* code,
                  std::string* err) {
  std::vector<const char*> names, texts;
  names.push_back("hip/hip_runtime.h");
  texts.push_back(kStubRuntime);
  names.push_back("math.h");
  texts.push_back(kStubMath);
  names.push_back("stdint.h");
  texts.push_back(kStubStdint);
  for (int i = 0; i < kJitHeaderCount; ++i) {
    names.push_back(kJitHeaderNames[i]);
    texts.push_back(kJitHeaderSources[i]);
  }
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "gpum_jit.hip", (int)names.size(),
                          texts.data(), names.data()) != HIPRTC_SUCCESS) {
    *err = "hiprtc: cannot create the program";
    return false;
  }
  // The library's own code-generation flags (build.py).
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-ffp-contract=on", "-Xclang", "-target-feature",
                        "-Xclang", "-packed-fp32-ops"};
  hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof *opts), opts);
  if (r != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    *err = std::string("hiprtc compile failed: ") + hiprtcGetErrorString(r) +
           "\n" + log.substr(0, 2000);
    return false;
  }
  size_t n = 0;
  hiprtcGetCodeSize(prog, &n);
  code->resize(n);
  hiprtcGetCode(prog, code->data());
  hiprtcDestroyProgram(&prog);
  return true;
}

bool compile(const std::string& src, Entry* out, std::string* err,
             double* ms) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<char> code;
  if (!compile_code(src, &code, err)) return false;
  hipError_t e = hipModuleLoadData(&out->mod, code.data());
  if (e == hipSuccess) e = hipModuleGetFunction(&out->fn, out->mod, "gpum_jit_kernel");
  if (e != hipSuccess) {
    *err = std::string("load compiled kernel: kernel/device error (") +
           hipGetErrorString(e) + ")";
    return false;
  }
  if (ms)
    *ms = std::chrono::duration<double, std::milli>(
              std::chrono::steady_clock::now() - t0).count();
  return true;
}

// hiprtc writes code objects for the ROCm release this library is built
// against; a process whose HIP runtime came from another release (a
// framework that bundles its own runtime, loaded first) may reject them, so
// the loaded runtime must be the release the library was compiled with.
bool runtime_matches(std::string* why) {
  int rt = 0;
  if (hipRuntimeGetVersion(&rt) != hipSuccess) {
    *why = "cannot query the HIP runtime version";
    return false;
  }
  const int rt_major = rt / 10000000, rt_minor = (rt / 100000) % 100;
  if (rt_major != HIP_VERSION_MAJOR || rt_minor != HIP_VERSION_MINOR) {
    std::ostringstream o;
    o << "compiled kernels need the HIP runtime this library was built with ("
      << HIP_VERSION_MAJOR << "." << HIP_VERSION_MINOR << "); runtime "
      << rt_major << "." << rt_minor << " is loaded in this process (e.g. "
         "bundled by another library imported first); use the interpreter";
    *why = o.str();
    return false;
  }
  return true;
}

// An FP64 constant written exactly (any value, -0.0 and NaN included).
std::string f64_literal(double v) {
  uint64_t bits;
  memcpy(&bits, &v, sizeof bits);
  std::ostringstream o;
  o << "__builtin_bit_cast(double, 0x" << std::hex << bits << "ull)";
  return o.str();
}

// A reference program over the parameters (hoist or prior) as straight-line
// FP64 statements: the operations of eval_params64 (step_common.h), one per
// statement, in its order, so the results are its results.  Returns the
// expression of the value left at the bottom of the stack ("0.0" if none).
std::string fp64_program(const uint32_t* code, size_t n_words,
                         const double* consts, std::ostringstream& o) {
  std::vector<std::string> st;
  int n = 0;
  auto fresh = [&]() { return "v" + std::to_string(n++); };
  for (size_t pc = 0; pc + 1 < n_words; pc += 2) {
    const uint32_t op = code[pc], arg = code[pc + 1];
    if (op == OP_PUSH_CONST) {
      st.push_back(f64_literal(consts[arg]));
    } else if (op == OP_PUSH_PARAM) {
      st.push_back("p[" + std::to_string(arg) + "]");
    } else if (op == OP_STORE_U || op == OP_STORE_U_RCP || op == OP_STORE_U_NEG) {
      const std::string a = st.back();
      st.pop_back();
      o << "    U[" << arg << "] = (float)("
        << (op == OP_STORE_U ? a : op == OP_STORE_U_RCP ? "1.0 / " + a : "-" + a)
        << ");\n";
    } else if (op == OP_SQUARE) {
      const std::string a = st.back();
      st.pop_back();
      const std::string v = fresh();
      o << "    const double " << v << " = " << a << " * " << a << ";\n";
      st.push_back(v);
    } else if (op == OP_NEG || op == OP_EXP || op == OP_LOG || op == OP_SQRT ||
               op == OP_LGAMMA) {
      const std::string a = st.back();
      st.pop_back();
      const std::string v = fresh();
      const char* f = op == OP_NEG ? "-" : op == OP_EXP ? "exp" : op == OP_LOG ? "log"
                      : op == OP_SQRT ? "sqrt" : "lgamma64";
      o << "    const double " << v << " = " << f << "(" << a << ");\n";
      st.push_back(v);
    } else {
      const std::string b = st.back();
      st.pop_back();
      const std::string a = st.back();
      st.pop_back();
      const std::string v = fresh();
      o << "    const double " << v << " = ";
      switch (op) {
        case OP_ADD: o << a << " + " << b; break;
        case OP_SUB: o << a << " - " << b; break;
        case OP_MUL: o << a << " * " << b; break;
        case OP_DIV: o << a << " / " << b; break;
        default: o << "pow(" << a << ", opaque64(" << b << "))"; break;
      }
      o << ";\n";
      st.push_back(v);
    }
  }
  return st.empty() ? "0.0" : st.front();
}

}  // namespace

std::string jit_eval64_source(const uint32_t* code, size_t n_words,
                              const double* consts) {
  const size_t n = n_words / 2;
  std::vector<uint32_t> lhs(n, 0), rhs(n, 0), st;
  // hoist[k]: an operation whose operands depend on no data (constants,
  // parameters, other such operations): computed once per point.
  // obs_only[k]: a data load or an operation that depends on no parameter:
  // computed once per observation for a group of points (pointwise).
  std::vector<char> data_dep(n, 0), hoist(n, 0), param_dep(n, 0), obs_only(n, 0);
  auto v = [](size_t k) { return "v" + std::to_string(k); };
  auto adj = [](size_t k) { return "a" + std::to_string(k); };
  auto is_unary = [](uint32_t op) {
    return op == OP_NEG || op == OP_EXP || op == OP_LOG || op == OP_SQRT ||
           op == OP_LGAMMA;
  };
  for (size_t pc = 0; pc < n; ++pc) {
    const uint32_t op = code[2 * pc];
    if (op == OP_PUSH_DATA) {
      data_dep[pc] = 1;
    } else if (op == OP_PUSH_PARAM) {
      param_dep[pc] = 1;
      hoist[pc] = 1;  // read once per point, held with the hoisted values
    } else if (op > OP_PUSH_DATA) {
      if (is_unary(op)) {
        lhs[pc] = st.back();
        st.pop_back();
        data_dep[pc] = data_dep[lhs[pc]];
        param_dep[pc] = param_dep[lhs[pc]];
      } else {
        rhs[pc] = st.back();
        st.pop_back();
        lhs[pc] = st.back();
        st.pop_back();
        data_dep[pc] = data_dep[lhs[pc]] | data_dep[rhs[pc]];
        param_dep[pc] = param_dep[lhs[pc]] | param_dep[rhs[pc]];
      }
      hoist[pc] = !data_dep[pc];
    }
    obs_only[pc] = data_dep[pc] && !param_dep[pc];
    st.push_back((uint32_t)pc);
  }
  // The value of instruction pc as an expression of its operands' names.
  auto expr = [&](size_t pc) -> std::string {
    const uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    const std::string l = v(lhs[pc]), r = v(rhs[pc]);
    switch (op) {
      case OP_PUSH_CONST: return f64_literal(consts[arg]);
      case OP_PUSH_PARAM: return "P[" + std::to_string(arg) + "]";
      case OP_PUSH_DATA:
        return "a.data[(size_t)" + std::to_string(arg) + " * a.n_obs + i]";
      case OP_ADD: return l + " + " + r;
      case OP_SUB: return l + " - " + r;
      case OP_MUL: return l + " * " + r;
      case OP_DIV: return l + " / " + r;
      case OP_NEG: return "-(" + l + ")";
      case OP_EXP: return "exp(" + l + ")";
      case OP_LOG: return "log(" + l + ")";
      case OP_SQRT: return "sqrt(" + l + ")";
      case OP_LGAMMA: return "lgamma64(" + l + ")";
      default: return "pow_ref64(" + l + ", " + r + ")";
    }
  };
  // Once per point: the hoisted values and the parameter-only factors of
  // their derivatives (digamma, and pow / log of a power's operands).
  std::ostringstream fields, prep;
  for (size_t pc = 0; pc < n; ++pc) {
    const uint32_t op = code[2 * pc];
    if (op == OP_PUSH_DATA) continue;
    if (op <= OP_PUSH_PARAM || hoist[pc])
      prep << "    const double " << v(pc) << " = " << expr(pc) << ";\n";
    if (!hoist[pc]) continue;
    fields << " double " << v(pc) << ";";
    prep << "    h." << v(pc) << " = " << v(pc) << ";\n";
    const std::string l = v(lhs[pc]), r = v(rhs[pc]);
    if (op == OP_LGAMMA) {
      fields << " double dg" << pc << ";";
      prep << "    h.dg" << pc << " = digamma64(" << l << ");\n";
    } else if (op == OP_POW) {
      fields << " double pw" << pc << "; double lg" << pc << ";";
      prep << "    h.pw" << pc << " = pow_ref64(" << l << ", " << r
           << " - 1.0);\n    h.lg" << pc << " = log(" << l << ");\n";
    }
  }
  std::ostringstream fwd, fwd_slice;
  for (size_t pc = 0; pc < n; ++pc)
    fwd << "    const double " << v(pc) << " = "
        << (hoist[pc] ? "h." + v(pc) : expr(pc)) << ";\n";
  // The same forward pass for an observation index below 2^29 (grad_fast
  // gets the index within a slice): a column's base plus a 32-bit byte
  // offset, which the load instruction takes as it is.
  for (size_t pc = 0; pc < n; ++pc) {
    fwd_slice << "    const double " << v(pc) << " = ";
    if (hoist[pc])
      fwd_slice << "h." << v(pc);
    else if (code[2 * pc] == OP_PUSH_DATA)
      fwd_slice << "*(const double*)((const char*)(a.data + (size_t)"
                << code[2 * pc + 1] << " * a.n_obs) + (i << 3))";
    else
      fwd_slice << expr(pc);
    fwd_slice << ";\n";
  }
  // Pointwise over a group of points: the observation's parameter-free
  // values once (obs), then the rest per point (value_o) -- the same
  // operations as `value`, so the same bits.
  std::ostringstream ofields, oprep, fwd_o;
  bool any_obs = false;
  for (size_t pc = 0; pc < n; ++pc) {
    const uint32_t op = code[2 * pc];
    if (op == OP_PUSH_CONST)
      oprep << "    const double " << v(pc) << " = " << expr(pc) << ";\n";
    if (!obs_only[pc]) continue;
    any_obs = true;
    ofields << " double " << v(pc) << ";";
    oprep << "    const double " << v(pc) << " = " << expr(pc) << ";\n"
          << "    o." << v(pc) << " = " << v(pc) << ";\n";
  }
  if (!any_obs) ofields << " double unused;";
  for (size_t pc = 0; pc < n; ++pc)
    fwd_o << "    const double " << v(pc) << " = "
          << (hoist[pc] ? "h." + v(pc) : obs_only[pc] ? "o." + v(pc) : expr(pc))
          << ";\n";
  const std::string root = n ? v(n - 1) : "0.0";
  // Backward: each value has one consumer, which defines its adjoint; a
  // zero adjoint propagates zeros (cpu_native.rs:218-277).  `bwd` applies
  // that rule with a select per adjoint.  `bwd_fast` leaves it out: every
  // adjoint is linear in its consumer's, so where the rule would give +0
  // the plain expression gives +0, -0 or NaN, all through the subtree; a
  // signed zero added to a gradient sum (which starts at +0) leaves its
  // bits unchanged, so the fast form equals the exact one bit for bit
  // unless a gradient sum ends NaN, and then the caller repeats the work
  // with `bwd`.
  std::ostringstream bwd, bwd_fast;
  if (n) {
    bwd << "    const double " << adj(n - 1) << " = 1.0;\n";
    bwd_fast << "    const double " << adj(n - 1) << " = 1.0;\n";
  }
  for (size_t k = n; k-- > 0;) {
    const uint32_t op = code[2 * k], arg = code[2 * k + 1];
    const std::string g = adj(k), live = "(" + adj(k) + " != 0.0)";
    const std::string l = v(lhs[k]), r = v(rhs[k]), me = v(k);
    const std::string kk = std::to_string(k);
    auto def = [&](uint32_t t, const std::string& e) {
      bwd << "    const double " << adj(t) << " = " << live << " ? (" << e
          << ") : 0.0;\n";
      bwd_fast << "    const double " << adj(t) << " = " << e << ";\n";
    };
    switch (op) {
      case OP_PUSH_CONST:
      case OP_PUSH_DATA: break;
      case OP_PUSH_PARAM:
        bwd << "    if " << live << " g[" << arg << "] += " << g << ";\n";
        bwd_fast << "    g[" << arg << "] += " << g << ";\n";
        break;
      case OP_ADD: def(lhs[k], g); def(rhs[k], g); break;
      case OP_SUB: def(lhs[k], g); def(rhs[k], "-" + g); break;
      case OP_MUL: def(lhs[k], g + " * " + r); def(rhs[k], g + " * " + l); break;
      case OP_DIV:
        def(lhs[k], g + " / " + r);
        def(rhs[k], "-(" + g + " * " + me + " / " + r + ")");
        break;
      case OP_NEG: def(lhs[k], "-" + g); break;
      case OP_EXP: def(lhs[k], g + " * " + me); break;
      case OP_LOG: def(lhs[k], g + " / " + l); break;
      case OP_SQRT: def(lhs[k], g + " * 0.5 / " + me); break;
      case OP_LGAMMA:
        def(lhs[k], g + " * " + (hoist[k] ? "h.dg" + kk : "digamma64(" + l + ")"));
        break;
      default: {
        const std::string pw =
            hoist[k] ? "h.pw" + kk : "pow_ref64(" + l + ", " + r + " - 1.0)";
        const std::string lg = hoist[k] ? "h.lg" + kk : "log(" + l + ")";
        def(lhs[k], g + " * " + r + " * " + pw);
        def(rhs[k], l + " > 0.0 ? " + g + " * " + me + " * " + lg + " : 0.0");
        break;
      }
    }
  }
  std::ostringstream o;
  o << "#include \"eval64_kernels.h\"\n#include \"step_common.h\"\n"
    << "namespace gpum {\n"
    << "struct JitEval64 {\n"
    << "  struct Hoisted {" << fields.str() << " };\n"
    << "  static __device__ __forceinline__ Hoisted prepare(const Eval64Args& a,\n"
    << "      const double* __restrict__ P) {\n"
    << "#pragma clang fp contract(off)\n"
    << "    Hoisted h;\n" << prep.str() << "    return h;\n  }\n"
    << "  static __device__ __forceinline__ double value(const Eval64Args& a,\n"
    << "      const Hoisted& h, const double* __restrict__ P, uint32_t i) {\n"
    << "#pragma clang fp contract(off)\n"
    << fwd.str() << "    return " << root << ";\n  }\n"
    << "  static __device__ __forceinline__ double grad(const Eval64Args& a,\n"
    << "      const Hoisted& h, const double* __restrict__ P, uint32_t i,\n"
    << "      double (&g)[kMaxParams]) {\n"
    << "#pragma clang fp contract(off)\n"
    << fwd.str() << bwd.str() << "    return " << root << ";\n  }\n"
    << "  static constexpr bool kFastGrad = true;\n"
    << "  static __device__ __forceinline__ double grad_fast(const Eval64Args& a,\n"
    << "      const Hoisted& h, const double* __restrict__ P, uint32_t i,\n"
    << "      double (&g)[kMaxParams]) {\n"
    << "#pragma clang fp contract(off)\n"
    << fwd_slice.str() << bwd_fast.str() << "    return " << root << ";\n  }\n"
    << "  struct ObsVals {" << ofields.str() << " };\n"
    << "  static __device__ __forceinline__ ObsVals obs(const Eval64Args& a,\n"
    << "      uint32_t i) {\n"
    << "#pragma clang fp contract(off)\n"
    << "    ObsVals o;\n" << oprep.str() << "    return o;\n  }\n"
    << "  static __device__ __forceinline__ double value_o(const Eval64Args& a,\n"
    << "      const Hoisted& h, const ObsVals& o, const double* __restrict__ P) {\n"
    << "#pragma clang fp contract(off)\n"
    << fwd_o.str() << "    return " << root << ";\n  }\n"
    << "};\n}  // namespace gpum\n";
  return o.str();
}

uint32_t unit_chunks_for(int ns, int opt, uint32_t n_cols) {
  const size_t chunk = 64u * (size_t)opt;
  return ns == 0 && 2 * chunk * std::max(1u, n_cols) * 4 <= (size_t)kTileF4 * 16
             ? 2u
             : 1u;
}

std::string jit_evaluator_source(const Lowered& L, const uint32_t* prior,
                                 size_t n_prior, const double* prior_consts,
                                 int ns, int opt, int cw) {
  std::ostringstream o;
  const int nsd = ns > 0 ? ns : 1;
  const uint32_t n_cols =
      L.empty ? 0u : (uint32_t)(L.n_cols_in + L.extra_cols.size());
  // Programs without spill slots whose uniforms fit a register budget keep
  // the wave's chains' uniform table in registers for a whole data pass
  // (read once per pass, not once per tile); the program then indexes that
  // copy with constant offsets.
  const int nup = std::max(1, L.n_uniform);
  const bool held = ns == 0 && L.sum_mode != SUM_AFFINE && nup * cw <= 64;
  o << "#include \"chain_body.h\"\n"
    << "namespace gpum {\n"
    << "struct JitEval {\n"
    << "  static constexpr int kOpt = " << opt << ";\n"
    << "  static constexpr int kCw = " << cw << ";\n"
    << "  static constexpr bool kPairs = " << (ns == 0 ? "true" : "false")
    << ";\n"
    << "  static constexpr int kUnit = " << unit_chunks_for(ns, opt, n_cols)
    << ";\n"
    << "  static constexpr bool kScalarCopy = true;\n";
  if (held)
    o << "  struct AccState { float u[" << cw * nup << "]; };\n";
  else
    o << "  struct AccState { float u[" << std::max<size_t>(1, L.vm.size())
      << "][" << cw << "]; };\n";
  o << "  static __device__ __forceinline__ void run(const ChainHot& H,\n"
    << "      const AccState& st, const float* __restrict__ U,\n"
    << "      const float* __restrict__ tile,\n"
    << "      uint32_t ld, int lane, float (&T)[" << cw << "][" << opt << "]) {\n"
    << "    float S[" << cw << "][" << nsd << "][" << opt << "];\n"
    << "    float D[" << opt << "];\n";
  if (held)
    o << "    const float* Ur = st.u;\n"
      << "    constexpr uint32_t nu = " << nup << "u;\n";
  else
    o << "    const float* Ur = U;\n"
      << "    const uint32_t nu = H.n_uniform_pad;\n";
  for (const VmInstr& ins : L.vm)
    o << "    vm_step<" << ns << ", " << opt << ", " << cw << ">(" << vm_op(ins)
      << "u, " << vm_slot(ins) << "u, " << vm_a(ins) << "u, " << vm_b(ins)
      << "u, Ur, nu, tile, ld, lane, T, S, D);\n";
  o << "  }\n"
    << "  static __device__ __forceinline__ uint32_t mode(const ChainHot&) {\n"
    << "    return " << L.sum_mode << "u;\n  }\n"
    << "  static __device__ __forceinline__ AccState acc_state(const ChainHot& H,\n"
    << "      const float* __restrict__ U) {\n"
    << "    AccState st;\n"
    << "    const uint32_t nu = H.n_uniform_pad;\n";
  // The uniforms of the affine form's products, or the whole table of the
  // wave's chains, read once per data pass.
  if (held)
    o << "    for (int j = 0; j < " << cw << "; ++j)\n"
      << "      for (int k = 0; k < " << nup << "; ++k)\n"
      << "        st.u[j * " << nup << " + k] = U[j * nu + k];\n";
  else if (L.sum_mode == SUM_AFFINE)
    for (size_t pc = 1; pc < L.vm.size(); ++pc)
      o << "    for (int j = 0; j < " << cw << "; ++j) st.u[" << pc
        << "][j] = U[j * nu + " << vm_a(L.vm[pc]) << "u];\n";
  o << "    return st;\n  }\n"
    << "  static __device__ __forceinline__ void run_acc(const ChainHot& H,\n"
    << "      const AccState& st, const float* __restrict__ U,\n"
    << "      const float* __restrict__ tile, uint32_t ld, int lane,\n"
    << "      float (&s)[" << cw << "]) {\n";
  // The interpreter's vm_run_acc order: instructions 1.. of the program,
  // added to the caller's sums.
  if (L.sum_mode == SUM_AFFINE)
    for (size_t pc = 1; pc < L.vm.size(); ++pc)
      o << "    vm_acc_fma_u<" << opt << ", " << cw << ">(st.u[" << pc
        << "], tile, ld, lane, " << vm_b(L.vm[pc]) << "u, s);\n";
  o << "  }\n"
    << "  static __device__ __forceinline__ double finish(const ChainHot& H,\n"
    << "      const float* U, double s, bool with_const) {\n";
  if (L.sum_tail.empty()) {
    o << "    return s;\n";
  } else {
    o << "    double alpha = 1.0, beta = 0.0;\n";
    for (const VmInstr& t : L.sum_tail)
      o << "    vm_tail_step(" << vm_op(t) << "u, (double)U[" << vm_a(t)
        << "], (double)U[" << vm_b(t) << "], alpha, beta);\n";
    o << "    return vm_tail_apply(alpha, beta, s, H.n_obs, with_const);\n";
  }
  // The callers' program lengths still decide: a launch that passes an
  // empty program (the data term alone, without the prior) gets what the
  // interpreter returns for one.
  o << "  }\n"
    << "  using Params = JitEval;\n"
    << "  static __device__ __forceinline__ void hoist(const uint32_t*, uint32_t n,\n"
    << "      const double*, const double* __restrict__ p, float* U) {\n"
    << "    if (n < 2) return;\n";
  fp64_program(L.hoist.data(), L.hoist.size(), L.consts.data(), o);
  o << "  }\n"
    << "  static __device__ __forceinline__ double prior(const uint32_t*, uint32_t n,\n"
    << "      const double*, const double* __restrict__ p) {\n"
    << "    if (n < 2) return 0.0;\n";
  const std::string r = fp64_program(prior, n_prior, prior_consts, o);
  o << "    return " << r << ";\n  }\n};\n}  // namespace gpum\n";
  return o.str();
}

hipFunction_t jit_kernel(const std::string& evaluator, JitKind kind,
                         std::string* err, double* compile_ms) {
  const std::string src = evaluator + kKernels[kind];
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_cache.find(src);
  if (it != g_cache.end()) {
    if (compile_ms) *compile_ms = 0.0;
    return it->second.fn;
  }
  if (!runtime_matches(err)) return nullptr;
  Entry e;
  if (!compile(src, &e, err, compile_ms)) return nullptr;
  int spill = 0;
  if (kind == JIT_GRAD64 &&
      hipFuncGetAttribute(&spill, HIP_FUNC_ATTRIBUTE_LOCAL_SIZE_BYTES, e.fn) ==
          hipSuccess &&
      spill > kGradSpillBytes) {
    Entry e2;
    double ms2 = 0.0;
    std::string err2;
    if (compile(evaluator + kGrad64TwoWaves, &e2, &err2, &ms2)) {
      hipModuleUnload(e.mod);
      e = e2;
      if (compile_ms) *compile_ms += ms2;
    }
  }
  g_cache[src] = e;
  return e.fn;
}

bool jit_compile_only(const std::string& evaluator, JitKind kind,
                      std::string* err, double* compile_ms) {
  auto t0 = std::chrono::steady_clock::now();
  std::vector<char> code;
  if (!compile_code(evaluator + kKernels[kind], &code, err)) return false;
  if (compile_ms)
    *compile_ms = std::chrono::duration<double, std::milli>(
                      std::chrono::steady_clock::now() - t0).count();
  return true;
}

}  // namespace gpum
