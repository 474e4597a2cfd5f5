// runtime.h -- internals of the host runtime shared by its source files.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpum_hip.h"
#include "chain_kernel.h"
#include "eval64.h"
#include "lower.h"

namespace gpum {
namespace rt {

void set_err(const char* fmt, ...);
// A HIP failure as a message carrying the phrases the R guard matches.
void set_hip_err(const char* what, hipError_t e);

#define HIP_TRY(what, expr)            \
  do {                                 \
    hipError_t e_ = (expr);            \
    if (e_ != hipSuccess) {            \
      ::gpum::rt::set_hip_err(what, e_); \
      return false;                    \
    }                                  \
  } while (0)

struct Device {
  bool ready = false;
  bool ok = false;
  int id = 0;
  int n_cu = 0;
  size_t lds_max = 0;
  hipStream_t stream = nullptr;
  std::string why;
};

extern std::mutex g_mu;  // serialises device work of the library
extern Device g_dev;
bool device_init();

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  ~DevBuf() { release(); }
  void release() {
    if (p) hipFree(p);
    p = nullptr;
    n = 0;
  }
  bool alloc(size_t count, const char* what) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      set_hip_err(what, e);
      return false;
    }
    n = count;
    return true;
  }
  bool upload(const T* src, size_t count, const char* what) {
    if (!alloc(count, what)) return false;
    if (count == 0 || src == nullptr) return true;
    hipError_t e = hipMemcpy(p, src, count * sizeof(T), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      set_hip_err(what, e);
      return false;
    }
    return true;
  }
};

// Kernel variant: spill slots, observations per lane per chunk, chains per
// wave.
struct Variant {
  int ns, opt, cw;
};

// The variant of a program of depth vm_depth summed in sum_mode (SUM_*).
Variant pick_variant(int vm_depth, uint32_t sum_mode);
uint32_t round_up(uint32_t x, uint32_t m);

struct Tiling {
  Variant v;
  // unit: chunks per FP32 partial (ChainHot::unit_chunks).
  uint32_t cpb, wpc, chunk, unit, tile_len, n_pad, resident, stream_mode,
      lds_tile_offset;
  size_t lds_bytes;
};

}  // namespace rt
}  // namespace gpum

struct gpum_hip_model {
  gpum::Lowered L;
  uint32_t n_params = 0;
  uint32_t n_cols_total = 0;
  uint32_t n_obs = 0;
  // host copies needed to lay the data out per tiling
  std::vector<float> cols;  // column-major, n_obs per column, unpadded
  // device program
  gpum::rt::DevBuf<uint32_t> vm, hoist, prior;
  gpum::rt::DevBuf<double> consts, prior_consts;
  gpum::rt::DevBuf<float> u_init;
  uint32_t n_uniform_pad = 1;
  std::vector<uint32_t> prior_code;
  std::vector<double> prior_consts_h;
  // device data, laid out for `data_ld`
  gpum::rt::DevBuf<float> data;
  uint32_t data_ld = 0;
  uint32_t data_chunk = 0;  // layout granularity the data was padded to
  // Launches run per-model compiled kernels (else the interpreter).
  bool jit = false;
  // The programs as given and the original column count, for the FP64
  // value-and-gradient path of Langevin proposals; its device state (tapes,
  // the FP32 columns widened to FP64) is built on first use.
  std::vector<uint32_t> code_ref, prior_ref;
  std::vector<double> consts_ref;
  uint32_t n_cols_in = 0;
  bool eval64_ready = false;
  uint32_t slots_ll = 0;  // values the gradient keeps for its backward pass
  uint32_t slots_prior = 0;
  uint32_t prior_depth = 0;
  gpum::rt::DevBuf<gpum::TapeOp> tape_ll, tape_prior;
  gpum::rt::DevBuf<double> data64, consts64;
};

namespace gpum {
namespace rt {

// Constraints a batched launch puts on each model's tiling.
struct TilingForce {
  Variant v;            // VM variant shared by the launch
  size_t extra_static;  // static LDS beyond the two tile buffers
};

// Tiling of a launch over n_chains chains (or, with `points`, of the points
// kernel over n_chains slots: one wave per `cw` slots, 16 waves per
// workgroup).  With `force`: the given variant, always streamed, room for
// extra static LDS.  `cw_req`: chains per wave asked for (0: automatic;
// ext->chains_per_wave takes its place when set).
bool pick_tiling(const gpum_hip_model* m, uint32_t n_chains,
                 const gpum_run_ext* ext, Tiling* t, bool points = false,
                 const TilingForce* force = nullptr, int cw_req = 0);
// Lay the model's columns out on the device with column stride n_pad.
bool ensure_data(gpum_hip_model* m, uint32_t n_pad);
// Langevin (MALA, proposal_mode 3) chains of gpum_hip_run_chains_ex /
// gpum_hip_model_run, run by the stepwise sampler; the library lock held.
int run_langevin(gpum_hip_model* m, const double* init, uint32_t n_chains,
                 const double* proposal_sd, const double* temperatures,
                 uint32_t n_iter, uint32_t seed, const double* proposal_l,
                 gpum_run_ext* ext, gpum_chain_result* out);
// The FP64 tapes and data of a model (Langevin proposals), built once.
bool ensure_eval64(gpum_hip_model* m);
// Validate, lower and upload a model (precompute / fuse: -1 default, 0, 1).
bool model_build(gpum_hip_model* m, const uint32_t* code, size_t code_len,
                 const double* consts, size_t n_consts,
                 const uint32_t* prior_code, size_t prior_len,
                 const double* prior_consts, size_t n_prior_consts,
                 uint32_t n_params, const double* data, uint32_t n_cols,
                 uint32_t n_obs, int32_t precompute, int32_t fuse);

// The summation fields of a launch's hot values (the tail is uploaded
// after the program's padding instruction).
inline bool model_affine(const gpum_hip_model* m) {
  return !m->L.empty && m->L.sum_mode == SUM_AFFINE;
}

// Whether the interpreter has kernels of variant v: every variant with one
// or two chains per wave; four for affine programs (variant (0, 16, 4), the
// affine interpreter).  Other four-chain launches need a compiled kernel.
inline bool interp_has_variant(const Variant& v) {
  return v.cw <= 2 || (v.ns == 0 && v.opt == 16);
}

inline uint32_t model_sum_mode(const gpum_hip_model* m) {
  return m->L.empty ? (uint32_t)SUM_VALUE : m->L.sum_mode;
}

inline void set_sum_fields(const gpum_hip_model* m, ChainHot* h) {
  h->sum_mode = m->L.empty ? (uint32_t)SUM_VALUE : m->L.sum_mode;
  h->n_sum_tail = m->L.empty ? 0u : (uint32_t)m->L.sum_tail.size();
}

// The model's compiled kernel of `kind` for variant v when the model runs
// compiled kernels and the launch can use one (module kernels get at most
// 64 KB of dynamic LDS), else nullptr; sets *failed (and the error) when
// compilation fails.
hipFunction_t model_jit_kernel(const gpum_hip_model* m, int kind,
                               const Variant& v, size_t lds_bytes,
                               bool* failed);
// Source of the model's compiled evaluator for variant v.
std::string model_jit_source(const gpum_hip_model* m, const Variant& v);
// Waves of a workgroup of tiling t.
inline uint32_t tiling_waves(const Tiling& t) {
  return t.cpb / (uint32_t)t.v.cw * t.wpc;
}
hipError_t launch_module(hipFunction_t f, dim3 grid, dim3 block,
                         size_t lds_bytes, hipStream_t s, void** args);

}  // namespace rt
}  // namespace gpum
