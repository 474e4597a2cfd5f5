// gpum_hip.cpp -- host runtime behind the C ABI in include/gpum_hip.h.
//
// The HIP replacement of the reference's CubeCL host launcher
// (src/rust/src/chain_kernel.rs:371-532): validate and lower the bytecode,
// lay the data out for the device (FP32, column-major, padded per tile,
// precomputed data-only columns appended), pick the tiling, launch the fused
// chain kernel on the library's stream, and return the reference's result
// layout.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/gpum_hip.h"
#include "chain_kernel.h"
#include "jit.h"
#include "lower.h"
#include "runtime.h"

using namespace gpum;
using namespace gpum::rt;

namespace {
thread_local std::string g_err;
}  // namespace

namespace gpum {
namespace rt {

void set_err(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
}

// A HIP failure as a message carrying the phrases the R guard matches.
void set_hip_err(const char* what, hipError_t e) {
  if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
    set_err("%s: out of memory on the device (%s)", what, hipGetErrorString(e));
  else if (e == hipErrorNoDevice || e == hipErrorInvalidDevice)
    set_err("%s: no usable device (%s)", what, hipGetErrorString(e));
  else
    set_err("%s: kernel/device error (%s)", what, hipGetErrorString(e));
}

std::mutex g_mu;
Device g_dev;

bool device_init() {
  if (g_dev.ready) {
    if (!g_dev.ok) set_err("%s", g_dev.why.c_str());
    return g_dev.ok;
  }
  g_dev.ready = true;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
    g_dev.why = "no HIP device available";
    set_err("%s", g_dev.why.c_str());
    return false;
  }
  int id = 0;
  hipGetDevice(&id);
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, id) != hipSuccess) {
    g_dev.why = "cannot query the HIP device";
    set_err("%s", g_dev.why.c_str());
    return false;
  }
  if (strncmp(p.gcnArchName, "gfx950", 6) != 0) {
    g_dev.why = std::string("device ") + p.gcnArchName +
                " is not gfx950 (MI355X); this backend is built for gfx950 only";
    set_err("%s", g_dev.why.c_str());
    return false;
  }
  g_dev.id = id;
  g_dev.n_cu = p.multiProcessorCount;
  g_dev.lds_max = p.sharedMemPerBlock;
  if (g_dev.lds_max < 65536) g_dev.lds_max = 65536;
  if (g_dev.lds_max > 160 * 1024) g_dev.lds_max = 160 * 1024;
  hipError_t le = set_chain_kernel_lds_limit(g_dev.lds_max);
  if (le != hipSuccess) {
    g_dev.why = std::string("cannot raise the kernels' LDS limit to ") +
                std::to_string(g_dev.lds_max) + " bytes: kernel/device error (" +
                hipGetErrorString(le) + ")";
    set_err("%s", g_dev.why.c_str());
    return false;
  }
  if (hipStreamCreateWithFlags(&g_dev.stream, hipStreamNonBlocking) != hipSuccess) {
    g_dev.why = "cannot create a HIP stream on the device";
    set_err("%s", g_dev.why.c_str());
    return false;
  }
  g_dev.ok = true;
  return true;
}

Variant pick_variant(int vm_depth, uint32_t sum_mode) {
  int need = vm_depth > 1 ? vm_depth - 1 : 0;
  // A depth-1 program that keeps a value per observation (not affine)
  // works on fewer observations per lane, so that several chains per wave
  // hold their values, their data and their uniforms: 8 for a value sum,
  // 4 for a square sum (the batched regression kernel spills 128 B per
  // lane at 8 with two chains per wave, none at 4).  The choice depends on
  // the program only, so every chains-per-wave count computes alike.
  if (need == 0)
    return {0, sum_mode == SUM_AFFINE ? 16 : sum_mode == SUM_SQUARE ? 4 : 8, 1};
  if (need <= 1) return {1, 16, 1};
  if (need <= 3) return {3, 16, 1};
  if (need <= 7) return {7, 8, 1};
  if (need <= 15) return {15, 4, 1};
  return {31, 4, 1};
}

uint32_t round_up(uint32_t x, uint32_t m) { return (x + m - 1) / m * m; }

}  // namespace rt
}  // namespace gpum

namespace gpum {
namespace rt {

bool model_build(gpum_hip_model* m, const uint32_t* code, size_t code_len,
                 const double* consts, size_t n_consts,
                 const uint32_t* prior_code, size_t prior_len,
                 const double* prior_consts, size_t n_prior_consts,
                 uint32_t n_params, const double* data, uint32_t n_cols,
                 uint32_t n_obs, int32_t precompute, int32_t fuse) {
  if (n_params == 0 || n_params > (uint32_t)kMaxParams) {
    set_err("n_params must be in [1, %d] on the hip backend (got %u)",
            kMaxParams, n_params);
    return false;
  }
  if (n_cols > 0 && n_obs > 0 && data == nullptr) {
    set_err("data is NULL but n_obs * n_cols > 0");
    return false;
  }
  if (n_consts > 0 && consts == nullptr) {
    set_err("consts is NULL");
    return false;
  }
  if (code_len > 0 && code == nullptr) {
    set_err("code is NULL");
    return false;
  }
  int pdepth = 0;
  std::string e = validate_program(prior_code, prior_len, n_prior_consts,
                                   n_params, 0, false, &pdepth);
  if (!e.empty()) {
    set_err("prior: %s", e.c_str());
    return false;
  }
  LowerOptions opt;
  opt.precompute_data = precompute != 0;
  opt.fuse = fuse != 0;
  e = lower_program(code, code_len, consts, n_consts, n_params, n_cols, opt,
                    &m->L);
  if (!e.empty()) {
    set_err("log-likelihood: %s", e.c_str());
    return false;
  }
  if (m->L.n_uniform > kMaxUniform) {
    set_err("too many uniform operands");
    return false;
  }
  m->n_params = n_params;
  m->n_obs = n_obs;
  m->code_ref.assign(code, code + code_len);
  m->consts_ref.assign(consts, consts + n_consts);
  m->prior_ref.assign(prior_code, prior_code + prior_len);
  m->prior_depth = (uint32_t)pdepth;
  m->n_cols_in = n_cols;
  // Columns: originals cast to FP32 (as the reference's GPU path casts,
  // chain_kernel.rs:501-505), then precomputed data-only subtrees evaluated
  // in FP64 on the original f64 values and rounded once.
  const uint32_t nx = (uint32_t)m->L.extra_cols.size();
  m->n_cols_total = n_cols + nx;
  if (m->L.empty) m->n_cols_total = 0;
  m->cols.assign((size_t)m->n_cols_total * n_obs, 0.0f);
  if (!m->L.empty) {
    for (uint32_t i = 0; i < n_obs; ++i) {
      const double* row = data + (size_t)i * n_cols;
      for (uint32_t c = 0; c < n_cols; ++c)
        m->cols[(size_t)c * n_obs + i] = (float)row[c];
      for (uint32_t k = 0; k < nx; ++k) {
        const std::vector<uint32_t>& pc = m->L.extra_cols[k].code;
        double v = eval_ref64(pc.data(), pc.size(), m->L.consts.data(),
                              nullptr, row);
        m->cols[(size_t)(n_cols + k) * n_obs + i] = (float)v;
      }
    }
  }
  // Device program.
  std::vector<uint32_t> vm4;
  for (const VmInstr& ins : m->L.vm) {
    vm4.push_back(vm_op(ins));
    vm4.push_back(vm_slot(ins));
    vm4.push_back(vm_a(ins));
    vm4.push_back(vm_b(ins));
  }
  // One padding instruction: the kernel prefetches pc + 1.
  for (int k = 0; k < 4; ++k) vm4.push_back(k == 0 ? (uint32_t)VM_NUM_OPS : 0u);
  // The sum tail follows the padding.
  for (const VmInstr& ins : m->L.sum_tail) {
    vm4.push_back(vm_op(ins));
    vm4.push_back(vm_slot(ins));
    vm4.push_back(vm_a(ins));
    vm4.push_back(vm_b(ins));
  }
  m->n_uniform_pad = (uint32_t)std::max(1, m->L.n_uniform);
  std::vector<float> u0(m->n_uniform_pad, 0.0f);
  for (auto& kv : m->L.const_uniforms) u0[kv.first] = kv.second;
  // The device's copy of the prior squares with OP_SQUARE (square_pow2).
  m->prior_code = square_pow2(prior_code, prior_len, prior_consts);
  m->prior_consts_h.assign(prior_consts, prior_consts + n_prior_consts);
  if (!m->vm.upload(vm4.data(), vm4.size(), "upload program")) return false;
  if (!m->hoist.upload(m->L.hoist.data(), m->L.hoist.size(), "upload program"))
    return false;
  if (!m->consts.upload(m->L.consts.data(), m->L.consts.size(), "upload consts"))
    return false;
  if (!m->prior.upload(m->prior_code.data(), m->prior_code.size(),
                       "upload prior"))
    return false;
  if (!m->prior_consts.upload(prior_consts, n_prior_consts, "upload prior"))
    return false;
  if (!m->u_init.upload(u0.data(), u0.size(), "upload uniforms")) return false;
  return true;
}

}  // namespace rt
}  // namespace gpum

namespace gpum {
namespace rt {

static bool pick_tiling_cw(const gpum_hip_model* m, uint32_t n_chains,
                           const gpum_run_ext* ext, Tiling* t, bool points,
                           int cw_in, const TilingForce* force);

bool pick_tiling(const gpum_hip_model* m, uint32_t n_chains,
                 const gpum_run_ext* ext, Tiling* t, bool points,
                 const TilingForce* force, int cw_req) {
  if (!force) {
    if (ext && ext->struct_size >= offsetof(gpum_run_ext, chains_per_wave) + 4 &&
        ext->chains_per_wave > 0)
      cw_req = ext->chains_per_wave;
    // Automatic: several chains per wave when the program keeps no stack,
    // the chains fill every CU with such groups and the launch runs a
    // compiled kernel or the affine interpreter (the general interpreter's
    // dispatch holds every chain's registers at each opcode merge and
    // spills).  A ladder or a pinned workgroup must hold whole groups.
    if (cw_req == 0) {
      const bool has_tiling =
          ext && ext->struct_size >= offsetof(gpum_run_ext, waves_per_chain) + 4;
      // Four chains per wave for affine programs on the chain kernel: each
      // LDS read feeds four chains' FMAs (4096 gamma chains over 1e6
      // observations: 67.7 ms per 250-iteration launch, against 71.9 with
      // two and 83.7 with one).  Otherwise two.
      for (uint32_t cw : {4u, 2u}) {
        if (cw == 4 && (points || !model_affine(m))) continue;
        const bool split_ladder = ext && ext->swap_every > 0 &&
                                  ext->swap_group > 1 && ext->swap_group % cw;
        const bool split_pin = has_tiling && ext->chains_per_block > 0 &&
                               (uint32_t)ext->chains_per_block % cw;
        if ((m->jit || model_affine(m)) && m->L.vm_depth <= 1 &&
            n_chains >= cw * (uint32_t)g_dev.n_cu && !split_ladder &&
            !split_pin &&
            pick_tiling_cw(m, n_chains, ext, t, points, (int)cw, nullptr) &&
            (!m->jit || t->lds_bytes <= kJitMaxDynamicLds))
          return true;
      }
      cw_req = 1;
    }
  }
  return pick_tiling_cw(m, n_chains, ext, t, points, force ? 0 : cw_req, force);
}

// pick_tiling with the chains per wave fixed (cw_in; with `force`, the
// forced variant's).
static bool pick_tiling_cw(const gpum_hip_model* m, uint32_t n_chains,
                           const gpum_run_ext* ext, Tiling* t, bool points,
                           int cw_in, const TilingForce* force) {
  t->v = force ? force->v : pick_variant(m->L.vm_depth, model_sum_mode(m));
  const size_t static_lds =
      kStreamStaticLds + (force ? force->extra_static
                                : (points ? kPointsExtraLds : 0));
  const uint32_t opt = (uint32_t)t->v.opt;
  t->chunk = 64u * opt;
  // Two chunks per FP32 partial (one conversion and FP64 add per two
  // chunks) for programs without spill slots whose two chunks of rows fit
  // one tile buffer.  The choice depends on the model alone (its program
  // picks the variant), so every tiling sums alike.
  t->unit = unit_chunks_for(t->v.ns, t->v.opt, m->n_cols_total);
  const uint32_t ncu = (uint32_t)g_dev.n_cu;
  const bool has_tiling =
      ext && ext->struct_size >= offsetof(gpum_run_ext, waves_per_chain) + 4;
  const bool pinned = has_tiling && ext->chains_per_block > 0;
  const bool ladder = ext && ext->swap_every > 0 && ext->swap_group > 1;
  // Chains per wave: two chains share every observation a wave reads from
  // LDS (half the LDS traffic per evaluation).
  if (!force) {
    if (cw_in != 1 && !((cw_in == 2 || cw_in == 4) && t->v.ns == 0)) {
      set_err("chains_per_wave must be 1, or 2 or 4 for a program of depth 1");
      return false;
    }
    if (cw_in == 4 && !m->jit && !model_affine(m)) {
      set_err("four chains per wave run on compiled kernels "
              "(gpum_hip_model_set_jit) or for affine programs only");
      return false;
    }
    t->v.cw = cw_in;
  }
  const uint32_t cw = (uint32_t)t->v.cw;
  const uint32_t groups = (n_chains + cw - 1) / cw;  // waves' worth of chains
  uint32_t wpc, cpb;
  if (groups >= 16u * ncu) {
    wpc = 1;
  } else {
    uint32_t w = 16u * ncu / std::max(1u, groups);
    wpc = 1;
    while (wpc * 2 <= w && wpc * 2 <= 16) wpc *= 2;
  }
  // No more waves per chain group than there are chunks to share.
  uint32_t nchunks = std::max(1u, (m->n_obs + t->chunk - 1) / t->chunk);
  while (wpc > 1 && wpc > nchunks) wpc /= 2;
  auto fill = [&](uint32_t w) {  // chains of a full workgroup of w-wave groups
    return std::min(std::max(1u, 16u / w), std::max(1u, groups)) * cw;
  };
  cpb = fill(wpc);
  if (has_tiling) {
    if (ext->waves_per_chain > 0) wpc = (uint32_t)ext->waves_per_chain;
    if (ext->chains_per_block > 0) cpb = (uint32_t)ext->chains_per_block;
  }
  // Wide rows: fewer waves per chain so one unit per wave fits in LDS.
  {
    const uint32_t ncol0 = std::max(1u, m->n_cols_total);
    const size_t tile_cap = (size_t)kTileF4 * 16;
    while (wpc > 1 &&
           (size_t)t->chunk * t->unit * wpc * ncol0 * 4 > tile_cap) {
      wpc /= 2;
      if (!pinned) cpb = fill(wpc);
    }
  }
  // A temperature ladder lives inside one workgroup.
  if (ladder) {
    const uint32_t G = ext->swap_group;
    while (wpc > 1 && G / cw * wpc > 16) wpc /= 2;
    cpb = G * std::max(1u, (16u / wpc * cw) / G);
    while (cpb > G && cpb > n_chains) cpb -= G;
  }
  if (points) {
    // One wave per cw slots: the points kernel splits observations over
    // workgroups instead of waves.
    wpc = 1;
    cpb = std::min(16u * cw, groups * cw);
  }
  if (cpb == 0 || wpc == 0 || cpb % cw != 0 || cpb / cw * wpc > 16) {
    set_err("invalid tiling: chains_per_block must be a multiple of "
            "chains_per_wave and chains_per_block / chains_per_wave * "
            "waves_per_chain at most 16");
    return false;
  }
  t->cpb = cpb;
  t->wpc = wpc;
  const uint32_t ncol = std::max(1u, m->n_cols_total);
  auto head_bytes = [&](uint32_t c, uint32_t w) {
    if (points) return ((size_t)c * m->n_uniform_pad * 4 + 15) / 16 * 16;
    const size_t state_bytes = (size_t)c * (6 * m->n_params + 6) * 8;
    const size_t part_bytes = (size_t)c * w * 8;
    const size_t u_bytes = (size_t)c * m->n_uniform_pad * 4;
    return (state_bytes + part_bytes + u_bytes + 15) / 16 * 16;
  };
  // Per-chain records and uniform tables must leave room for the two
  // streamed tile buffers.
  while (!pinned && cpb > cw &&
         head_bytes(cpb, wpc) + static_lds > g_dev.lds_max) {
    cpb = std::max(1u, cpb / cw / 2) * cw;
  }
  t->cpb = cpb;
  size_t head = head_bytes(cpb, wpc);
  t->lds_tile_offset = (uint32_t)head;
  const uint32_t step = t->chunk * t->unit * wpc;  // one unit per wave
  const size_t tile_floats = (size_t)kTileF4 * 4;  // one static buffer
  uint32_t n_res = round_up(std::max(1u, m->n_obs), step);
  if (!force && (size_t)n_res * ncol * 4 <= (size_t)96 * 1024 &&
      head + (size_t)n_res * ncol * 4 + (points ? kPointsExtraLds : 0) <=
          g_dev.lds_max) {
    // Resident: the whole padded data set in dynamic LDS.
    t->resident = 1;
    t->stream_mode = 0;
    t->tile_len = n_res;
    t->n_pad = n_res;
    t->lds_bytes = head + (size_t)n_res * ncol * 4;
    return true;
  }
  if (head + static_lds > g_dev.lds_max) {
    set_err("per-chain state too large for the LDS budget (%u chains per "
            "block, %u parameters)", cpb, m->n_params);
    return false;
  }
  if ((size_t)step * ncol > tile_floats) {
    set_err("data rows too wide for the LDS tile (%u columns)", ncol);
    return false;
  }
  t->resident = 0;
  // The double-buffered copy is laid out for 1024-thread workgroups; other
  // shapes copy synchronously.
  t->stream_mode = cpb / cw * wpc == 16 ? 1 : 2;
  // Tiles of as many steps as one static buffer holds, an even number of
  // them (the double buffer alternates in pairs), as little padding as
  // that allows.
  uint32_t max_mult = (uint32_t)(tile_floats / ((size_t)step * ncol));
  uint32_t n_steps = (std::max(1u, m->n_obs) + step - 1) / step;
  uint32_t n_tiles = (n_steps + max_mult - 1) / max_mult;
  if (t->stream_mode == 1) n_tiles += n_tiles & 1u;
  uint32_t mult = (n_steps + n_tiles - 1) / n_tiles;
  t->tile_len = step * std::max(1u, mult);
  t->n_pad = t->tile_len * n_tiles;
  t->lds_bytes = head;
  return true;
}

std::string model_jit_source(const gpum_hip_model* m, const Variant& v) {
  return jit_evaluator_source(m->L, m->prior_code.data(), m->prior_code.size(),
                              m->prior_consts_h.data(), v.ns, v.opt, v.cw);
}

hipFunction_t model_jit_kernel(const gpum_hip_model* m, int kind,
                               const Variant& v, size_t lds_bytes,
                               bool* failed) {
  *failed = false;
  if (!m->jit || lds_bytes > kJitMaxDynamicLds) return nullptr;
  std::string err;
  hipFunction_t f = jit_kernel(model_jit_source(m, v),
                               (JitKind)kind, &err);
  if (!f) {
    set_err("%s", err.c_str());
    *failed = true;
  }
  return f;
}

hipError_t launch_module(hipFunction_t f, dim3 grid, dim3 block,
                         size_t lds_bytes, hipStream_t s, void** args) {
  return hipModuleLaunchKernel(f, grid.x, grid.y, grid.z, block.x, block.y,
                               block.z, (unsigned)lds_bytes, s, args, nullptr);
}

bool ensure_eval64(gpum_hip_model* m) {
  if (m->eval64_ready) return true;
  const std::vector<TapeOp> tl =
      make_tape(m->code_ref.data(), m->code_ref.size() / 2, &m->slots_ll);
  const std::vector<TapeOp> tp =
      make_tape(m->prior_ref.data(), m->prior_ref.size() / 2, &m->slots_prior);
  // The original columns as the kernels see them (FP32), widened.
  std::vector<double> d((size_t)m->n_cols_in * m->n_obs);
  if (!m->L.empty)
    for (size_t k = 0; k < d.size(); ++k) d[k] = (double)m->cols[k];
  if (!m->tape_ll.upload(tl.data(), tl.size(), "upload program") ||
      !m->tape_prior.upload(tp.data(), tp.size(), "upload prior") ||
      !m->consts64.upload(m->consts_ref.data(), m->consts_ref.size(),
                          "upload consts") ||
      !m->data64.upload(d.data(), d.size(), "upload data"))
    return false;
  m->eval64_ready = true;
  return true;
}

bool ensure_data(gpum_hip_model* m, uint32_t n_pad) {
  if (m->data.p && m->data_ld == n_pad) return true;
  const uint32_t ncol = std::max(1u, m->n_cols_total);
  std::vector<float> h((size_t)ncol * n_pad, 0.0f);
  for (uint32_t c = 0; c < m->n_cols_total; ++c)
    memcpy(&h[(size_t)c * n_pad], &m->cols[(size_t)c * m->n_obs],
           (size_t)m->n_obs * 4);
  if (!m->data.upload(h.data(), h.size(), "upload data")) return false;
  m->data_ld = n_pad;
  return true;
}

}  // namespace rt
}  // namespace gpum

namespace {

bool validate_run(const gpum_hip_model* m, const double* init,
                  uint32_t n_chains, const double* proposal_sd,
                  const double* temperatures, uint32_t proposal_mode,
                  const double* proposal_l) {
  if (n_chains == 0) {
    set_err("n_chains must be positive");
    return false;
  }
  if (!init || !proposal_sd || !temperatures) {
    set_err("init, proposal_sd and temperatures must not be NULL");
    return false;
  }
  if (proposal_mode == 3) {
    // Routed to the Langevin sampler before a fused launch is prepared.
    set_err("proposal_mode 3 (MALA) runs on the stepwise sampler");
    return false;
  }
  if (proposal_mode > 3) {
    set_err("unknown proposal_mode %u", proposal_mode);
    return false;
  }
  if (proposal_mode == 2 && proposal_l == nullptr) {
    set_err("proposal_mode 2 needs proposal_l");
    return false;
  }
  if (proposal_mode == 1 && n_chains < 2) {
    set_err("proposal_mode 1 (DE) needs at least 2 chains");
    return false;
  }
  for (uint32_t c = 0; c < n_chains; ++c) {
    if (!(temperatures[c] > 0.0) || !isfinite(temperatures[c])) {
      set_err("temperatures must be positive and finite");
      return false;
    }
  }
  (void)m;
  return true;
}

}  // namespace

extern "C" {

int gpum_hip_available(void) {
  std::lock_guard<std::mutex> lk(g_mu);
  return device_init() ? 1 : 0;
}

const char* gpum_hip_last_error(void) { return g_err.c_str(); }

void gpum_hip_result_free(gpum_chain_result* r) {
  if (!r) return;
  free(r->draws);
  free(r->accept_rate);
  free(r->last_log_post);
  r->draws = nullptr;
  r->accept_rate = nullptr;
  r->last_log_post = nullptr;
}

gpum_hip_model* gpum_hip_model_create(
    const uint32_t* code, size_t code_len, const double* consts,
    size_t n_consts, const uint32_t* prior_code, size_t prior_len,
    const double* prior_consts, size_t n_prior_consts, uint32_t n_params,
    const double* data, uint32_t n_cols, uint32_t n_obs,
    int32_t precompute_data, int32_t fuse) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return nullptr;
  hipSetDevice(g_dev.id);
  gpum_hip_model* m = new gpum_hip_model();
  if (!model_build(m, code, code_len, consts, n_consts, prior_code, prior_len,
                   prior_consts, n_prior_consts, n_params, data, n_cols, n_obs,
                   precompute_data, fuse)) {
    delete m;
    return nullptr;
  }
  return m;
}

int gpum_hip_model_set_jit(gpum_hip_model* m, int32_t enable,
                           double* compile_ms) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!m) {
    set_err("null model");
    return 1;
  }
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  if (compile_ms) *compile_ms = 0.0;
  if (!enable || m->L.empty) {
    m->jit = false;
    return 0;
  }
  // Compile the streamed chain kernel of the variant of large launches now
  // (a failure is reported here); other kinds and variants on first use.
  Variant v = pick_variant(m->L.vm_depth, model_sum_mode(m));
  if (v.ns == 0) v.cw = 2;
  std::string src = model_jit_source(m, v);
  std::string err;
  double ms = 0.0;
  if (!jit_kernel(src, JIT_CHAIN_STREAM, &err, &ms)) {
    set_err("%s", err.c_str());
    return 1;
  }
  m->jit = true;
  if (compile_ms) *compile_ms = ms;
  return 0;
}

void gpum_hip_model_free(gpum_hip_model* m) {
  std::lock_guard<std::mutex> lk(g_mu);
  delete m;
}

static void fill_info(const Lowered& L, uint32_t n_cols_total,
                      gpum_model_info* info) {
  info->n_vm_instr = (uint32_t)L.vm.size();
  info->vm_depth = (uint32_t)L.vm_depth;
  info->ref_depth = (uint32_t)L.ref_depth;
  info->n_uniform = (uint32_t)L.n_uniform;
  info->n_cols_total = n_cols_total;
  info->n_hoist_words = (uint32_t)L.hoist.size();
  info->w_exec = L.w_exec;
  info->w_bytecode = L.w_bytecode;
  info->sum_mode = L.empty ? 0u : L.sum_mode;
  info->n_sum_tail = L.empty ? 0u : (uint32_t)L.sum_tail.size();
}

int gpum_hip_model_info(const gpum_hip_model* m, gpum_model_info* info) {
  if (!m || !info) {
    set_err("null model or info");
    return 1;
  }
  fill_info(m->L, m->n_cols_total, info);
  return 0;
}

int gpum_hip_lower_info(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        uint32_t n_params, uint32_t n_cols,
                        int32_t precompute_data, int32_t fuse,
                        gpum_model_info* info) {
  if (!info || (code_len && !code) || (n_consts && !consts)) {
    set_err("null program, constants or info");
    return 1;
  }
  if (n_params > (uint32_t)kMaxParams) {
    set_err("at most %d parameters", kMaxParams);
    return 1;
  }
  LowerOptions opt;
  opt.precompute_data = precompute_data != 0;
  opt.fuse = fuse != 0;
  Lowered L;
  std::string e = lower_program(code, code_len, consts, n_consts, n_params,
                                n_cols, opt, &L);
  if (!e.empty()) {
    set_err("log-likelihood: %s", e.c_str());
    return 1;
  }
  if (L.n_uniform > kMaxUniform) {
    set_err("too many uniform operands");
    return 1;
  }
  fill_info(L, L.empty ? 0u : n_cols + (uint32_t)L.extra_cols.size(), info);
  return 0;
}

int gpum_hip_jit_check(const uint32_t* code, size_t code_len,
                       const double* consts, size_t n_consts,
                       const uint32_t* prior_code, size_t prior_len,
                       const double* prior_consts, size_t n_prior_consts,
                       uint32_t n_params, uint32_t n_cols,
                       int32_t chains_per_wave, uint32_t kinds,
                       double* compile_ms) {
  if ((code_len && !code) || (n_consts && !consts) ||
      (prior_len && !prior_code) || (n_prior_consts && !prior_consts)) {
    set_err("null program or constants");
    return 1;
  }
  if (n_params > (uint32_t)kMaxParams) {
    set_err("at most %d parameters", kMaxParams);
    return 1;
  }
  std::string e = validate_program(prior_code, prior_len, n_prior_consts,
                                   n_params, 0, false, nullptr);
  if (!e.empty()) {
    set_err("prior: %s", e.c_str());
    return 1;
  }
  LowerOptions opt;
  Lowered L;
  e = lower_program(code, code_len, consts, n_consts, n_params, n_cols, opt, &L);
  if (!e.empty()) {
    set_err("log-likelihood: %s", e.c_str());
    return 1;
  }
  Variant v = pick_variant(L.vm_depth, L.empty ? (uint32_t)SUM_VALUE : L.sum_mode);
  const int cw = chains_per_wave > 0 ? chains_per_wave : 1;
  if (cw != 1 && !((cw == 2 || cw == 4) && v.ns == 0)) {
    set_err("chains_per_wave must be 1, or 2 or 4 for a program of depth 1");
    return 1;
  }
  v.cw = cw;
  const std::vector<uint32_t> prior = square_pow2(prior_code, prior_len,
                                                  prior_consts);
  const std::string src = jit_evaluator_source(L, prior.data(), prior.size(),
                                               prior_consts, v.ns, v.opt, v.cw);
  const std::string src64 = jit_eval64_source(code, code_len, consts);
  double total = 0.0;
  for (int k = 0; k < JIT_NUM_KINDS; ++k) {
    if (!(kinds & (1u << k))) continue;
    std::string err;
    double ms = 0.0;
    if (!jit_compile_only(k < kJitChainKinds ? src : src64, (JitKind)k, &err,
                          &ms)) {
      set_err("%s", err.c_str());
      return 1;
    }
    total += ms;
  }
  if (compile_ms) *compile_ms = total;
  return 0;
}

static int model_run_locked(gpum_hip_model* m, const double* init,
                            uint32_t n_chains, const double* proposal_sd,
                            const double* temperatures, uint32_t n_iter,
                            uint32_t seed, uint32_t proposal_mode,
                            double gamma, double de_noise,
                            const double* proposal_l, gpum_run_ext* ext,
                            gpum_chain_result* out, double* loglik_out,
                            bool with_prior = true) {
  if (!out && !loglik_out) {
    set_err("null result");
    return 1;
  }
  // Fields past the caller's struct_size read as zero (ABI growth).
  gpum_run_ext X;
  memset(&X, 0, sizeof X);
  gpum_run_ext* caller_ext = ext;
  if (ext) {
    memcpy(&X, ext, std::min<size_t>(ext->struct_size, sizeof X));
    ext = &X;
  }
  uint32_t n_batches = 0;
  if (ext && ext->swap_every > 0) {
    n_batches = std::max(1u, n_iter / ext->swap_every);
    if (ext->swap_group > 1) {
      if (ext->swap_group > 16 || n_chains % ext->swap_group != 0) {
        set_err("swap_group must divide n_chains and be at most 16 (got %u "
                "for %u chains)", ext->swap_group, n_chains);
        return 1;
      }
    }
    if (ext->adapt_batches > n_batches) ext->adapt_batches = n_batches;
  }
  if (!validate_run(m, init, n_chains, proposal_sd, temperatures,
                    proposal_mode, proposal_l))
    return 1;
  hipSetDevice(g_dev.id);
  const uint32_t np = m->n_params;
  Tiling t;
  if (!pick_tiling(m, n_chains, ext, &t)) return 1;
  if (!ensure_data(m, t.n_pad)) return 1;

  const size_t ncp = (size_t)n_chains * np;
  const size_t n_draws = (size_t)n_iter * ncp;
  DevBuf<double> d_init, d_sd, d_temp, d_l, d_lp, d_final, d_tplp, d_tprop, d_tinit;
  DevBuf<double> d_iscale, d_fsd, d_fscale, d_tll;
  DevBuf<float> d_draws, d_acc, d_acch;
  DevBuf<uint8_t> d_tacc, d_swaph;
  DevBuf<ChainArgs> d_args;
  auto fail = [&]() { return 1; };
  if (!d_init.upload(init, ncp, "upload init")) return fail();
  if (!d_sd.upload(proposal_sd, ncp, "upload proposal_sd")) return fail();
  if (!d_temp.upload(temperatures, n_chains, "upload temperatures")) return fail();
  if (proposal_mode == 2 &&
      !d_l.upload(proposal_l, ncp * np, "upload proposal_l"))
    return fail();
  if (!d_draws.alloc(n_draws, "allocate draws")) return fail();
  if (!d_acc.alloc(n_chains, "allocate accept_rate")) return fail();
  if (!d_lp.alloc(n_chains, "allocate last_log_post")) return fail();
  const bool want_final = ext && ext->final_state;
  const bool want_tplp = ext && ext->trace_plp;
  const bool want_tprop = ext && ext->trace_prop;
  const bool want_tacc = ext && ext->trace_accept;
  const bool want_tinit = ext && ext->init_log_post;
  const bool want_tll = ext && ext->trace_ll;
  if (want_final && !d_final.alloc(ncp, "allocate final_state")) return fail();
  if (want_tplp && !d_tplp.alloc((size_t)n_iter * n_chains, "allocate trace")) return fail();
  if (want_tprop && !d_tprop.alloc(n_draws, "allocate trace")) return fail();
  if (want_tacc && !d_tacc.alloc((size_t)n_iter * n_chains, "allocate trace")) return fail();
  if (want_tinit && !d_tinit.alloc(n_chains, "allocate trace")) return fail();
  if (want_tll && !d_tll.alloc((size_t)n_iter * n_chains, "allocate trace")) return fail();
  const bool want_swaph = ext && ext->swap_hist && n_batches > 0;
  const bool want_acch = ext && ext->acc_hist && n_batches > 0;
  const bool want_fsd = ext && ext->final_sd;
  const bool want_fscale = ext && ext->final_scale;
  if (ext && ext->init_scale &&
      !d_iscale.upload(ext->init_scale, n_chains, "upload init_scale"))
    return fail();
  if (want_swaph && !d_swaph.alloc((size_t)n_batches * n_chains, "allocate swap history"))
    return fail();
  if (want_swaph && hipMemset(d_swaph.p, 0xFF, (size_t)n_batches * n_chains) != hipSuccess) {
    set_err("clear swap history: device error");
    return fail();
  }
  if (want_acch && !d_acch.alloc((size_t)n_batches * n_chains, "allocate accept history"))
    return fail();
  if (want_fsd && !d_fsd.alloc(ncp, "allocate final_sd")) return fail();
  if (want_fscale && !d_fscale.alloc(n_chains, "allocate final_scale")) return fail();

  ChainArgs a;
  memset(&a, 0, sizeof a);
  a.vm = reinterpret_cast<const VmIns4*>(m->vm.p);
  a.n_vm = (uint32_t)m->L.vm.size();
  a.hoist = m->hoist.p;
  a.n_hoist = (uint32_t)m->L.hoist.size();
  a.consts = m->consts.p;
  a.prior = m->prior.p;
  a.n_prior = with_prior ? (uint32_t)m->prior_code.size() : 0u;
  a.prior_consts = m->prior_consts.p;
  a.u_init = m->u_init.p;
  a.n_uniform_pad = m->n_uniform_pad;
  a.data = m->data.p;
  a.n_cols = m->n_cols_total;
  a.n_obs = m->n_obs;
  a.n_pad = t.n_pad;
  a.ld = m->data_ld;
  a.n_chains = n_chains;
  a.n_params = np;
  a.init = d_init.p;
  a.proposal_sd = d_sd.p;
  a.temperatures = d_temp.p;
  a.proposal_l = proposal_mode == 2 ? d_l.p : nullptr;
  a.gamma = gamma;
  a.de_noise = de_noise;
  a.mode = proposal_mode;
  a.n_iter = n_iter;
  a.seed = seed;
  a.stream_id = ext ? ext->stream_id : 0;
  a.draws = d_draws.p;
  a.accept_rate = d_acc.p;
  a.last_log_post = d_lp.p;
  a.final_state = want_final ? d_final.p : nullptr;
  a.trace_plp = want_tplp ? d_tplp.p : nullptr;
  a.trace_prop = want_tprop ? d_tprop.p : nullptr;
  a.trace_accept = want_tacc ? d_tacc.p : nullptr;
  a.trace_ll = want_tll ? d_tll.p : nullptr;
  a.init_log_post = want_tinit ? d_tinit.p : nullptr;
  a.iter_offset = ext ? ext->iter_offset : 0;
  a.swap_every = ext ? ext->swap_every : 0;
  a.n_batches = n_batches;
  a.swap_group = ext ? ext->swap_group : 0;
  a.adapt_batches = ext ? ext->adapt_batches : 0;
  a.batch0 = ext ? ext->batch0 : 0;
  a.init_scale = (ext && ext->init_scale) ? d_iscale.p : nullptr;
  a.swap_hist = want_swaph ? d_swaph.p : nullptr;
  a.acc_hist = want_acch ? d_acch.p : nullptr;
  a.final_sd = want_fsd ? d_fsd.p : nullptr;
  a.final_scale = want_fscale ? d_fscale.p : nullptr;
  a.cpb = t.cpb;
  a.wpc = t.wpc;
  a.tile_len = t.tile_len;
  a.resident = t.resident;
  a.lds_tile_offset = t.lds_tile_offset;
  if (!d_args.upload(&a, 1, "upload launch arguments")) return fail();

  ChainHot h;
  memset(&h, 0, sizeof h);
  h.vm = a.vm;
  h.u_init = a.u_init;
  h.data = a.data;
  h.n_vm = m->L.empty ? 0 : a.n_vm;
  h.n_uniform_pad = a.n_uniform_pad;
  h.n_cols = a.n_cols;
  h.n_obs = a.n_obs;
  h.n_pad = a.n_pad;
  h.ld = a.ld;
  h.n_chains = n_chains;
  h.n_iter = n_iter;
  h.cpb = t.cpb;
  h.wpc = t.wpc;
  h.tile_len = t.tile_len;
  h.unit_chunks = t.unit;
  h.stream_mode = t.stream_mode;
  h.lds_tile_offset = t.lds_tile_offset;
  h.state_words = 6 * np + 6;
  h.swap_every = a.swap_every;
  h.n_batches = a.n_batches;
  h.swap_group = a.swap_group;
  set_sum_fields(m, &h);

  bool jit_failed = false;
  hipFunction_t jf = model_jit_kernel(
      m, t.stream_mode ? JIT_CHAIN_STREAM : JIT_CHAIN_RESIDENT, t.v,
      t.lds_bytes, &jit_failed);
  if (jit_failed) return 1;
  if (!jf && !interp_has_variant(t.v)) {
    set_err("%d chains per wave need a compiled kernel, which this launch "
            "(%zu B of LDS) cannot use", t.v.cw, (size_t)t.lds_bytes);
    return 1;
  }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, g_dev.stream);
  hipError_t e;
  if (jf) {
    const ChainArgs* ap = d_args.p;
    void* kargs[] = {(void*)&ap, (void*)&h};
    e = launch_module(jf, dim3((h.n_chains + h.cpb - 1) / h.cpb),
                      dim3(tiling_waves(t) * 64), t.lds_bytes, g_dev.stream,
                      kargs);
  } else {
    e = launch_chain_kernel(d_args.p, h, t.v.ns, t.v.opt, t.v.cw, t.lds_bytes,
                            g_dev.stream);
  }
  hipEventRecord(e1, g_dev.stream);
  if (e != hipSuccess) {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    set_hip_err("chain kernel launch", e);
    return 1;
  }
  e = hipStreamSynchronize(g_dev.stream);
  if (e != hipSuccess) {
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    set_hip_err("chain kernel", e);
    return 1;
  }
  float ms = 0.0f;
  hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (caller_ext &&
      caller_ext->struct_size >= offsetof(gpum_run_ext, kernel_ms) + sizeof(double))
    caller_ext->kernel_ms = ms;

  if (loglik_out) {
    e = hipMemcpy(loglik_out, d_lp.p, n_chains * sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
      set_hip_err("read results", e);
      return 1;
    }
    return 0;
  }
  out->n_iter = n_iter;
  out->n_chains = n_chains;
  out->n_params = np;
  out->draws = (float*)malloc(std::max<size_t>(1, n_draws) * sizeof(float));
  out->accept_rate = (float*)malloc(n_chains * sizeof(float));
  out->last_log_post = (double*)malloc(n_chains * sizeof(double));
  if (!out->draws || !out->accept_rate || !out->last_log_post) {
    gpum_hip_result_free(out);
    set_err("host out of memory for results");
    return 1;
  }
  bool ok = hipMemcpy(out->draws, d_draws.p, n_draws * sizeof(float),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(out->accept_rate, d_acc.p, n_chains * sizeof(float),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(out->last_log_post, d_lp.p, n_chains * sizeof(double),
                      hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_final)
    ok = hipMemcpy(ext->final_state, d_final.p, ncp * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_tplp)
    ok = hipMemcpy(ext->trace_plp, d_tplp.p, (size_t)n_iter * n_chains * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_tll)
    ok = hipMemcpy(ext->trace_ll, d_tll.p, (size_t)n_iter * n_chains * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_tprop)
    ok = hipMemcpy(ext->trace_prop, d_tprop.p, n_draws * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_tinit)
    ok = hipMemcpy(ext->init_log_post, d_tinit.p, n_chains * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_swaph)
    ok = hipMemcpy(ext->swap_hist, d_swaph.p, (size_t)n_batches * n_chains,
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_acch)
    ok = hipMemcpy(ext->acc_hist, d_acch.p, (size_t)n_batches * n_chains * sizeof(float),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_fsd)
    ok = hipMemcpy(ext->final_sd, d_fsd.p, ncp * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_fscale)
    ok = hipMemcpy(ext->final_scale, d_fscale.p, n_chains * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && want_tacc)
    ok = hipMemcpy(ext->trace_accept, d_tacc.p, (size_t)n_iter * n_chains,
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    gpum_hip_result_free(out);
    set_err("read results: device error");
    return 1;
  }
  return 0;
}

int gpum_hip_model_run(gpum_hip_model* m, const double* init,
                       uint32_t n_chains, const double* proposal_sd,
                       const double* temperatures, uint32_t n_iter,
                       uint32_t seed, uint32_t proposal_mode, double gamma,
                       double de_noise, const double* proposal_l,
                       gpum_run_ext* ext, gpum_chain_result* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!m) {
    set_err("null model");
    return 1;
  }
  if (!device_init()) return 1;
  if (proposal_mode == 3)
    return run_langevin(m, init, n_chains, proposal_sd, temperatures, n_iter,
                        seed, proposal_l, ext, out);
  return model_run_locked(m, init, n_chains, proposal_sd, temperatures, n_iter,
                          seed, proposal_mode, gamma, de_noise, proposal_l,
                          ext, out, nullptr);
}

int gpum_hip_model_loglik(gpum_hip_model* m, const double* points,
                          uint32_t n_points, double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!m || !out) {
    set_err("null model or output");
    return 1;
  }
  if (!device_init()) return 1;
  if (n_points == 0) return 0;
  // A zero-iteration run evaluates the data term at the initial states;
  // the prior is left out (the batch is the data term only).
  std::vector<double> sd((size_t)n_points * m->n_params, 1.0);
  std::vector<double> temps(n_points, 1.0);
  return model_run_locked(m, points, n_points, sd.data(), temps.data(), 0, 0,
                          0, 0.0, 0.0, nullptr, nullptr, nullptr, out, false);
}

int gpum_hip_run_chains_ex(const uint32_t* code, size_t code_len,
                           const double* consts, size_t n_consts,
                           const uint32_t* prior_code, size_t prior_len,
                           const double* prior_consts, size_t n_prior_consts,
                           uint32_t n_params, const double* data,
                           uint32_t n_cols, uint32_t n_obs, const double* init,
                           uint32_t n_chains, const double* proposal_sd,
                           const double* temperatures, uint32_t n_iter,
                           uint32_t seed, uint32_t proposal_mode, double gamma,
                           double de_noise, const double* proposal_l,
                           gpum_run_ext* ext, gpum_chain_result* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  int32_t pre = -1, fu = -1;
  if (ext && ext->struct_size >= offsetof(gpum_run_ext, fuse) + 4) {
    pre = ext->precompute_data;
    fu = ext->fuse;
  }
  gpum_hip_model m;
  if (!model_build(&m, code, code_len, consts, n_consts, prior_code, prior_len,
                   prior_consts, n_prior_consts, n_params, data, n_cols, n_obs,
                   pre, fu))
    return 1;
  if (proposal_mode == 3)
    return run_langevin(&m, init, n_chains, proposal_sd, temperatures, n_iter,
                        seed, proposal_l, ext, out);
  return model_run_locked(&m, init, n_chains, proposal_sd, temperatures,
                          n_iter, seed, proposal_mode, gamma, de_noise,
                          proposal_l, ext, out, nullptr);
}

int gpum_hip_run_chains(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        const uint32_t* prior_code, size_t prior_len,
                        const double* prior_consts, size_t n_prior_consts,
                        uint32_t n_params, const double* data, uint32_t n_cols,
                        uint32_t n_obs, const double* init, uint32_t n_chains,
                        const double* proposal_sd, const double* temperatures,
                        uint32_t n_iter, uint32_t seed, uint32_t proposal_mode,
                        double gamma, double de_noise,
                        const double* proposal_l, gpum_chain_result* out) {
  return gpum_hip_run_chains_ex(code, code_len, consts, n_consts, prior_code,
                                prior_len, prior_consts, n_prior_consts,
                                n_params, data, n_cols, n_obs, init, n_chains,
                                proposal_sd, temperatures, n_iter, seed,
                                proposal_mode, gamma, de_noise, proposal_l,
                                nullptr, out);
}

int gpum_hip_loglik_batch(const uint32_t* code, size_t code_len,
                          const double* consts, size_t n_consts,
                          uint32_t n_params, const double* data,
                          uint32_t n_cols, uint32_t n_obs,
                          const double* points, uint32_t n_points,
                          double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  if (n_points == 0) return 0;
  if (!points || !out) {
    set_err("points and out must not be NULL");
    return 1;
  }
  gpum_hip_model m;
  if (!model_build(&m, code, code_len, consts, n_consts, nullptr, 0, nullptr,
                   0, n_params, data, n_cols, n_obs, -1, -1))
    return 1;
  std::vector<double> sd((size_t)n_points * n_params, 1.0);
  std::vector<double> temps(n_points, 1.0);
  return model_run_locked(&m, points, n_points, sd.data(), temps.data(), 0, 0,
                          0, 0.0, 0.0, nullptr, nullptr, nullptr, out);
}

}  // extern "C"
