// warmup.cpp -- gpum_hip_model_warmup: the adaptive warmup of
// R/adaptive.R:101-254 with every batch launch and every between-batch
// step on the device, captured in one HIP graph.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/gpum_hip.h"
#include "chain_kernel.h"
#include "jit.h"
#include "runtime.h"
#include "warmup.h"

using namespace gpum;
using namespace gpum::rt;

namespace {

// Optimal random-walk acceptance by dimension (R/adaptive.R:69-72).
double target_accept_rw(uint32_t d) {
  static const double low_d[8] = {0.44, 0.35, 0.32, 0.30, 0.28, 0.27, 0.26, 0.25};
  return d >= 1 && d <= 8 ? low_d[d - 1] : 0.234;
}

// The batch plan of R/adaptive.R:112-118: 10 to 40 batches of about 50
// iterations, the last absorbing the remainder.
std::vector<uint32_t> warmup_plan(uint32_t warmup) {
  uint32_t nb = std::min(40u, std::max(10u, warmup / 50u));
  nb = std::min(nb, warmup);
  const uint32_t k = warmup / nb;
  std::vector<uint32_t> sizes(nb, k);
  sizes[nb - 1] += warmup - nb * k;
  return sizes;
}

}  // namespace

extern "C" int gpum_hip_model_warmup(gpum_hip_model* m, const double* init,
                                     uint32_t n_chains,
                                     const double* proposal_sd,
                                     gpum_warmup_opts* opts_in) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!m || !init || !proposal_sd || !opts_in) {
    set_err("warmup needs a model, init, proposal_sd and options");
    return 1;
  }
  gpum_warmup_opts o;
  memset(&o, 0, sizeof o);
  memcpy(&o, opts_in, std::min<size_t>(opts_in->struct_size, sizeof o));
  if (o.warmup == 0) {
    set_err("warmup needs warmup > 0");
    return 1;
  }
  if (n_chains == 0) {
    set_err("n_chains must be positive");
    return 1;
  }
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  const uint32_t np = m->n_params;
  const size_t ncp = (size_t)n_chains * np;
  for (size_t i = 0; i < ncp; ++i) {
    if (!std::isfinite(init[i])) {
      set_err("init must be finite");
      return 1;
    }
    if (!(proposal_sd[i] > 0.0) || !std::isfinite(proposal_sd[i])) {
      set_err("proposal_sd must be positive and finite");
      return 1;
    }
  }
  const std::vector<uint32_t> sizes = warmup_plan(o.warmup);
  const uint32_t nb = (uint32_t)sizes.size();
  const uint32_t reset = nb / 2;
  const uint32_t t_max = *std::max_element(sizes.begin(), sizes.end());
  // Robbins-Monro gains: the batch index restarts after the restart batch.
  std::vector<double> gain(nb);
  for (uint32_t b = 1; b <= nb; ++b)
    gain[b - 1] = pow((double)(b > reset ? b - reset : b), -2.0 / 3.0);
  // Batch b runs the Cholesky proposal once the pooled accumulator has held
  // two draws before it (the factor, once set, is kept through the
  // restart); batch 1 and one-parameter models run the diagonal one.
  std::vector<uint32_t> mode(nb, 0u);
  {
    double pool_n = 0.0;
    bool has_l = false;
    for (uint32_t b = 1; b <= nb; ++b) {
      mode[b - 1] = np > 1 && has_l ? 2u : 0u;
      pool_n += (double)n_chains * sizes[b - 1];
      if (b == reset) pool_n = 0.0;
      if (np > 1 && pool_n >= 2.0) has_l = true;
    }
  }

  gpum_run_ext ext;
  memset(&ext, 0, sizeof ext);
  ext.struct_size = sizeof ext;
  Tiling t;
  if (!pick_tiling(m, n_chains, &ext, &t)) return 1;
  if (!ensure_data(m, t.n_pad)) return 1;

  DevBuf<double> d_init, d_sd, d_sd0, d_temp, d_scale, d_prev, d_L, d_lp;
  DevBuf<double> d_wn, d_wmean, d_wm2, d_wsig, d_bmean, d_bm2;
  DevBuf<double> d_pn, d_pmean, d_pm2, d_psig, d_lpool, d_diag, d_gain;
  DevBuf<float> d_draws, d_acc, d_hist;
  DevBuf<uint32_t> d_ctl, d_sizes;
  DevBuf<ChainArgs> d_args;
  std::vector<double> ones(n_chains, 1.0);
  if (!d_init.upload(init, ncp, "upload init") ||
      !d_sd0.upload(proposal_sd, ncp, "upload proposal_sd") ||
      !d_sd.alloc(ncp, "allocate warmup state") ||
      !d_temp.upload(ones.data(), n_chains, "allocate warmup state") ||
      !d_scale.alloc(n_chains, "allocate warmup state") ||
      !d_prev.alloc(n_chains, "allocate warmup state") ||
      !d_L.alloc(ncp * np, "allocate warmup state") ||
      !d_lp.alloc(n_chains, "allocate warmup state") ||
      !d_wn.alloc(n_chains, "allocate warmup state") ||
      !d_wmean.alloc(ncp, "allocate warmup state") ||
      !d_wm2.alloc(ncp, "allocate warmup state") ||
      !d_wsig.alloc(ncp, "allocate warmup state") ||
      !d_bmean.alloc(ncp, "allocate warmup state") ||
      !d_bm2.alloc(ncp * np, "allocate warmup state") ||
      !d_pn.alloc(1, "allocate warmup state") ||
      !d_pmean.alloc(np, "allocate warmup state") ||
      !d_pm2.alloc((size_t)np * np, "allocate warmup state") ||
      !d_psig.alloc(np, "allocate warmup state") ||
      !d_lpool.alloc((size_t)np * np, "allocate warmup state") ||
      !d_diag.alloc(np, "allocate warmup state") ||
      !d_gain.upload(gain.data(), nb, "allocate warmup state") ||
      !d_draws.alloc((size_t)t_max * ncp, "allocate draws") ||
      !d_acc.alloc(n_chains, "allocate warmup state") ||
      !d_hist.alloc((size_t)nb * n_chains, "allocate warmup state") ||
      !d_ctl.alloc(kWarmCtlWords, "allocate warmup state") ||
      !d_sizes.upload(sizes.data(), nb, "allocate warmup state"))
    return 1;

  WarmupDev w;
  memset(&w, 0, sizeof w);
  w.n_chains = n_chains;
  w.np = np;
  w.n_batches = nb;
  w.reset_batch = reset;
  w.auto_stop = o.auto_stop;
  w.target = target_accept_rw(np);
  w.scale0 = 2.38 / sqrt((double)np);
  w.sizes = d_sizes.p;
  w.gain = d_gain.p;
  w.draws = d_draws.p;
  w.accept = d_acc.p;
  w.init = d_init.p;
  w.sd = d_sd.p;
  w.scale = d_scale.p;
  w.prev_scale = d_prev.p;
  w.L = d_L.p;
  w.w_n = d_wn.p;
  w.w_mean = d_wmean.p;
  w.w_m2 = d_wm2.p;
  w.w_sig = d_wsig.p;
  w.b_mean = d_bmean.p;
  w.b_m2 = d_bm2.p;
  w.p_n = d_pn.p;
  w.p_mean = d_pmean.p;
  w.p_m2 = d_pm2.p;
  w.p_sig = d_psig.p;
  w.L_pool = d_lpool.p;
  w.diag_now = d_diag.p;
  w.ctl = d_ctl.p;
  w.acc_hist = d_hist.p;

  // One argument block per batch: its seed, length and proposal mode.
  std::vector<ChainArgs> args(nb);
  for (uint32_t b = 1; b <= nb; ++b) {
    ChainArgs& a = args[b - 1];
    memset(&a, 0, sizeof a);
    a.vm = reinterpret_cast<const VmIns4*>(m->vm.p);
    a.n_vm = (uint32_t)m->L.vm.size();
    a.hoist = m->hoist.p;
    a.n_hoist = (uint32_t)m->L.hoist.size();
    a.consts = m->consts.p;
    a.prior = m->prior.p;
    a.n_prior = (uint32_t)m->prior_code.size();
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
    a.proposal_l = mode[b - 1] == 2 ? d_L.p : nullptr;
    a.mode = mode[b - 1];
    a.n_iter = sizes[b - 1];
    a.seed = o.seed + b - 1;
    a.draws = d_draws.p;
    a.accept_rate = d_acc.p;
    a.last_log_post = d_lp.p;
    a.stop = d_ctl.p + kWarmFrozenAt;
    a.cpb = t.cpb;
    a.wpc = t.wpc;
    a.tile_len = t.tile_len;
    a.resident = t.resident;
    a.lds_tile_offset = t.lds_tile_offset;
  }
  if (!d_args.upload(args.data(), nb, "upload launch arguments")) return 1;
  ChainHot h;
  memset(&h, 0, sizeof h);
  h.vm = args[0].vm;
  h.u_init = args[0].u_init;
  h.data = args[0].data;
  h.n_vm = m->L.empty ? 0 : args[0].n_vm;
  h.n_uniform_pad = args[0].n_uniform_pad;
  h.n_cols = args[0].n_cols;
  h.n_obs = args[0].n_obs;
  h.n_pad = t.n_pad;
  h.ld = args[0].ld;
  h.n_chains = n_chains;
  h.cpb = t.cpb;
  h.wpc = t.wpc;
  h.tile_len = t.tile_len;
  h.unit_chunks = t.unit;
  h.stream_mode = t.stream_mode;
  h.lds_tile_offset = t.lds_tile_offset;
  h.state_words = 6 * np + 6;
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
  const dim3 grid((n_chains + t.cpb - 1) / t.cpb);
  const dim3 block(tiling_waves(t) * 64);
  hipStream_t st = g_dev.stream;
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = launch_warmup_init(w, d_sd0.p, st);
    for (uint32_t b = 1; b <= nb && e == hipSuccess; ++b) {
      ChainHot hb = h;
      hb.n_iter = sizes[b - 1];
      const ChainArgs* ap = d_args.p + (b - 1);
      if (jf) {
        void* kargs[] = {(void*)&ap, (void*)&hb};
        e = launch_module(jf, grid, block, t.lds_bytes, st, kargs);
      } else {
        e = launch_chain_kernel(ap, hb, t.v.ns, t.v.opt, t.v.cw, t.lds_bytes,
                                st);
      }
      if (e == hipSuccess) e = launch_warmup_chains(w, b, sizes[b - 1], st);
      if (e == hipSuccess) e = launch_warmup_pool(w, b, sizes[b - 1], st);
      if (e == hipSuccess) e = launch_warmup_post(w, b, st);
    }
    return e;
  };
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipError_t e = hipSuccess;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  if (o.use_graph) {
    e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
      hipError_t e2 = enqueue();
      e = hipStreamEndCapture(st, &graph);
      if (e2 != hipSuccess) e = e2;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = hipGraphLaunch(exec, st);
    if (e == hipSuccess) e = hipEventRecord(e1, st);
  } else {
    e = hipEventRecord(e0, st);
    if (e == hipSuccess) e = enqueue();
    if (e == hipSuccess) e = hipEventRecord(e1, st);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0.0f;
  if (e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) {
    set_hip_err("device warmup", e);
    return 1;
  }
  uint32_t ctl[kWarmCtlWords];
  bool ok = hipMemcpy(ctl, d_ctl.p, sizeof ctl, hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && o.final_init)
    ok = hipMemcpy(o.final_init, d_init.p, ncp * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && o.final_sd)
    ok = hipMemcpy(o.final_sd, d_sd.p, ncp * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && o.final_scale)
    ok = hipMemcpy(o.final_scale, d_scale.p, n_chains * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && o.final_l)
    ok = hipMemcpy(o.final_l, d_L.p, ncp * np * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && o.accept_hist)
    ok = hipMemcpy(o.accept_hist, d_hist.p, (size_t)nb * n_chains * sizeof(float),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    set_err("read results: device error");
    return 1;
  }
  opts_in->n_batches_planned = nb;
  opts_in->n_batches = ctl[kWarmBatches];
  opts_in->warmup_used = ctl[kWarmUsed];
  opts_in->has_l = np > 1 ? ctl[kWarmHasL] : 0u;
  opts_in->kernel_ms = ms;
  return 0;
}
