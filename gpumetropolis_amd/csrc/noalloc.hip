// noalloc.hip -- the gradient batch without device allocation.
//
// gpum_hip_grad_batch_noalloc evaluates what gpum_hip_grad_batch does (the
// gradient of the summed data term at many parameter points, FP64 reverse
// mode) for callers that cannot call hipMalloc: code under HIP graph capture,
// tight serving loops, harnesses that forbid it.  Data, points and output are
// device buffers the caller owns; everything else -- the program's tape, its
// constants and the per-slice partial sums -- lives in the file-scope array
// `scratch` below, one copy per GPU for the life of the process.
//
// The array's name is a host-side stand-in on the host: its device address
// is looked up once per device with hipGetSymbolAddress (the device current)
// and only that address is used afterwards.  The kernels are the library's
// own interpreted FP64 kernels (eval64.hip): a per-program compiled kernel
// would load a code object, which allocates.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "eval64.h"
#include "runtime.h"

namespace gpum {
namespace {

// Scratch layout: [tape | constants | partial sums].
constexpr size_t kTapeBytes = (size_t)kMaxTapeOps * sizeof(TapeOp);
constexpr size_t kMaxScratchConsts = 4096;
constexpr size_t kConstsOffset = (kTapeBytes + 255) / 256 * 256;
constexpr size_t kPartialOffset =
    (kConstsOffset + kMaxScratchConsts * sizeof(double) + 255) / 256 * 256;
constexpr size_t kScratchBytes = size_t(8) << 20;
constexpr size_t kPartialBytes = kScratchBytes - kPartialOffset;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` per device id (looked up on first use).
constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};

bool scratch_ptr(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    rt::set_err("device id %d out of range for the no-allocation scratch", dev);
    return false;
  }
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      rt::set_hip_err("looking up the no-allocation scratch (device symbol)",
                      e == hipSuccess ? hipErrorInvalidSymbol : e);
      return false;
    }
    g_scratch[dev] = p;
  }
  *out = static_cast<unsigned char*>(g_scratch[dev]);
  return true;
}

}  // namespace
}  // namespace gpum

using namespace gpum;

extern "C" {

int gpum_hip_grad_batch_noalloc(const uint32_t* code, size_t code_len,
                                const double* consts, size_t n_consts,
                                uint32_t n_params, const double* d_data,
                                uint32_t n_cols, uint32_t n_obs,
                                const double* d_points, uint32_t n_points,
                                double* d_out) {
  std::lock_guard<std::mutex> lk(rt::g_mu);
  if (!rt::device_init()) return 1;
  hipSetDevice(rt::g_dev.id);
  if (code_len > 0 && !code) {
    rt::set_err("code must not be NULL");
    return 1;
  }
  int depth = 0;
  const std::string e0 = validate_program(code, code_len, n_consts, n_params,
                                          n_cols, true, &depth);
  if (!e0.empty()) {
    rt::set_err("%s", e0.c_str());
    return 1;
  }
  if (n_params == 0) n_points = 0;
  if (n_points == 0) return 0;
  if (!d_points || !d_out) {
    rt::set_err("points and out must not be NULL");
    return 1;
  }
  if (n_params > (uint32_t)kMaxParams) {
    rt::set_err("the gradient supports at most %d parameters", kMaxParams);
    return 1;
  }
  if (code_len / 2 > kMaxTapeOps) {
    rt::set_err("the gradient supports programs of at most %u instructions",
                kMaxTapeOps);
    return 1;
  }
  if (n_consts > kMaxScratchConsts) {
    rt::set_err("the no-allocation gradient takes at most %zu constants "
                "(got %zu)", kMaxScratchConsts, n_consts);
    return 1;
  }
  if (n_consts > 0 && !consts) {
    rt::set_err("consts must not be NULL");
    return 1;
  }
  hipStream_t st = rt::g_dev.stream;
  if (code_len == 0 || n_obs == 0) {
    // No term depends on the parameters: the gradient is zero.
    hipError_t e = hipMemsetAsync(d_out, 0,
                                  (size_t)n_points * n_params * sizeof(double),
                                  st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      rt::set_hip_err("clearing the gradient output", e);
      return 1;
    }
    return 0;
  }
  if (!d_data) {
    rt::set_err("data must not be NULL");
    return 1;
  }
  // One point's partials: every slice's gradient and value (64-bit: a
  // 32-bit n_obs + kGradSliceObs - 1 would wrap).
  const uint64_t ns64 = ((uint64_t)n_obs + kGradSliceObs - 1) / kGradSliceObs;
  const uint64_t per_point = ns64 * (n_params + 1) * sizeof(double);
  if (per_point > kPartialBytes) {
    rt::set_err("n_obs %u with %u parameters needs %zu bytes of partial sums "
                "per point; the no-allocation scratch holds %zu",
                n_obs, n_params, (size_t)per_point, kPartialBytes);
    return 1;
  }
  const uint32_t ns = (uint32_t)ns64;
  unsigned char* base = nullptr;
  if (!scratch_ptr(rt::g_dev.id, &base)) return 1;
  uint32_t n_slots = 0;
  const std::vector<TapeOp> tape = make_tape(code, code_len / 2, &n_slots);
  TapeOp* d_tape = reinterpret_cast<TapeOp*>(base);
  double* d_consts = reinterpret_cast<double*>(base + kConstsOffset);
  double* d_part = reinterpret_cast<double*>(base + kPartialOffset);
  // The tape and constants are copied into the scratch; the copies read
  // host memory, so they finish before this call returns.
  hipError_t e = hipMemcpyAsync(d_tape, tape.data(),
                                tape.size() * sizeof(TapeOp),
                                hipMemcpyHostToDevice, st);
  if (e == hipSuccess && n_consts > 0)
    e = hipMemcpyAsync(d_consts, consts, n_consts * sizeof(double),
                       hipMemcpyHostToDevice, st);
  Eval64Args a;
  memset(&a, 0, sizeof a);
  a.prog = d_tape;
  a.n_ops = (uint32_t)(code_len / 2);
  a.n_params = n_params;
  a.consts = d_consts;
  a.data = d_data;
  a.n_obs = n_obs;
  a.depth = (uint32_t)depth;
  a.n_slots = n_slots;
  a.n_slices = ns;
  a.out = d_part;
  const uint32_t chunk =
      (uint32_t)std::max<size_t>(1, std::min<size_t>(n_points, kPartialBytes / per_point));
  const uint64_t cap = (uint64_t)rt::g_dev.n_cu * 8u;
  // Chunks run in stream order: each reuses the partials of the one before.
  for (uint32_t p0 = 0; e == hipSuccess && p0 < n_points; p0 += chunk) {
    const uint32_t np = std::min(chunk, n_points - p0);
    a.points = d_points + (size_t)p0 * n_params;
    a.n_points = np;
    const uint32_t grid =
        (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(grad64_items(ns, np), cap));
    e = launch_grad64(a, grid, st);
    if (e == hipSuccess)
      e = launch_grad64_finish(d_part, np, n_params, ns,
                               d_out + (size_t)p0 * n_params, nullptr, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    rt::set_hip_err("gradient batch (no allocation)", e);
    return 1;
  }
  return 0;
}

}  // extern "C"
