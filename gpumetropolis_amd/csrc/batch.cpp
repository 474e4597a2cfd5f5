// batch.cpp -- many independent models, every batch of iterations, in one
// persistent launch.
//
// The reference runs one model per call (R/api.R:247-599); a workload of
// hundreds of models is a host loop of calls, each uploading its data
// (src/rust/src/chain_kernel.rs:393-413) and each batch of iterations
// another call (R/adaptive.R:427-524).  Here every model's program and data
// stay on the device, and one launch of the batched chain kernel advances
// every chain of every model through every batch: a workgroup takes work
// items (model, chain group, batch) from its XCD's queue, an item starts as
// soon as its own model has ended the previous batch, and the workgroup
// that ends a model's batch runs the reference's between-batch step for it.
// A model's items sit in one XCD's queue (they share its data in that L2),
// each queue batch-major and longest model first, so no batch waits for the
// slowest model of the previous one.
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "../../include/gpum_hip.h"
#include "batch.h"
#include "chain_kernel.h"
#include "jit.h"
#include "runtime.h"

using namespace gpum;
using namespace gpum::rt;

namespace {

struct ModelPlan {
  Tiling t;
  uint32_t chain_off;
  size_t draws_off;
};

}  // namespace

extern "C" int gpum_hip_batch_run(gpum_batch_model* models, uint32_t n_models,
                                  gpum_batch_opts* opts_in) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  if (!models || n_models == 0 || !opts_in) {
    set_err("batch needs at least one model and options");
    return 1;
  }
  gpum_batch_opts o;
  memset(&o, 0, sizeof o);
  memcpy(&o, opts_in, std::min<size_t>(opts_in->struct_size, sizeof o));
  if (o.proposal_mode != 0 && o.proposal_mode != 1) {
    set_err("batched runs take proposal_mode 0 or 1 (got %u)", o.proposal_mode);
    return 1;
  }
  if (o.n_iter == 0) {
    set_err("n_iter must be positive");
    return 1;
  }
  const uint32_t every = o.batch_len ? o.batch_len : o.n_iter;
  const uint32_t nb = std::max(1u, o.n_iter / every);
  std::vector<uint32_t> sizes(nb, every);
  sizes[nb - 1] += o.n_iter - nb * every;
  const uint32_t cw = o.chains_per_wave > 0 ? (uint32_t)o.chains_per_wave : 1u;
  const uint32_t cpb =
      o.chains_per_block > 0 ? (uint32_t)o.chains_per_block : 16u * cw;
  const uint32_t wpc = o.waves_per_chain > 0 ? (uint32_t)o.waves_per_chain
                                             : 16u / std::max(1u, cpb / cw);
  if ((cw != 1 && cw != 2 && cw != 4) || cpb % cw != 0 ||
      cpb / cw * wpc != 16) {
    set_err("batched runs need chains_per_wave 1, 2 or 4, chains_per_block "
            "a multiple of it and chains_per_block / chains_per_wave * "
            "waves_per_chain == 16");
    return 1;
  }
  int max_depth = 0;
  bool all_affine = true, all_square = true;
  for (uint32_t k = 0; k < n_models; ++k) {
    const gpum_batch_model& b = models[k];
    if (!b.model || !b.init || !b.proposal_sd || b.n_chains == 0) {
      set_err("model %u: model, init and proposal_sd must be set and "
              "n_chains positive", k);
      return 1;
    }
    if (o.proposal_mode == 1 && b.n_chains < 2) {
      set_err("model %u: proposal_mode 1 (DE) needs at least 2 chains", k);
      return 1;
    }
    max_depth = std::max(max_depth, b.model->L.vm_depth);
    all_affine = all_affine && model_affine(b.model);
    all_square = all_square && model_sum_mode(b.model) == SUM_SQUARE;
  }
  TilingForce force;
  force.v = pick_variant(max_depth, all_affine   ? (uint32_t)SUM_AFFINE
                                    : all_square ? (uint32_t)SUM_SQUARE
                                                 : (uint32_t)SUM_VALUE);
  if (cw > 1 && force.v.ns != 0) {
    set_err("chains_per_wave %u needs every model's program at depth 1", cw);
    return 1;
  }
  force.v.cw = (int)cw;
  force.extra_static = kBatchArgsLds;
  gpum_run_ext ext;
  memset(&ext, 0, sizeof ext);
  ext.struct_size = sizeof ext;
  ext.chains_per_block = (int32_t)cpb;
  ext.waves_per_chain = (int32_t)wpc;
  std::vector<ModelPlan> plan(n_models);
  size_t lds = 0;
  uint32_t n_chains = 0;
  size_t n_draws = 0, n_state = 0, n_welford = 0;
  for (uint32_t k = 0; k < n_models; ++k) {
    gpum_hip_model* m = models[k].model;
    if (!pick_tiling(m, models[k].n_chains, &ext, &plan[k].t, false, &force))
      return 1;
    if (plan[k].t.stream_mode != 1) {
      set_err("model %u: tiling is not streamable in a batched launch", k);
      return 1;
    }
    if (!ensure_data(m, plan[k].t.n_pad)) return 1;
    lds = std::max(lds, plan[k].t.lds_bytes);
    plan[k].chain_off = n_chains;
    plan[k].draws_off = n_draws;
    n_chains += models[k].n_chains;
    n_draws += (size_t)o.n_iter * models[k].n_chains * m->n_params;
    n_state += (size_t)models[k].n_chains * m->n_params;
    n_welford += (size_t)models[k].n_chains * (1 + 2 * m->n_params);
  }
  // Flat per-chain buffers; model k's chains are a contiguous slice.
  std::vector<double> h_snap, h_sd, h_temps(n_chains, 1.0);
  std::vector<size_t> state_off(n_models), welford_off(n_models);
  size_t so = 0, wo = 0;
  for (uint32_t k = 0; k < n_models; ++k) {
    const uint32_t C = models[k].n_chains, np = models[k].model->n_params;
    state_off[k] = so;
    welford_off[k] = wo;
    h_snap.insert(h_snap.end(), models[k].init, models[k].init + (size_t)C * np);
    h_sd.insert(h_sd.end(), models[k].proposal_sd,
                models[k].proposal_sd + (size_t)C * np);
    so += (size_t)C * np;
    wo += (size_t)C * (1 + 2 * np);
  }
  DevBuf<double> snap, sd, sd0, welford, temps, lp;
  DevBuf<float> draws, accept, acc_hist;
  DevBuf<uint32_t> sync;
  DevBuf<ChainArgs> d_args;
  DevBuf<ChainHot> d_hots;
  DevBuf<BatchModelInfo> d_info;
  DevBuf<BatchWork> d_work;
  DevBuf<BatchQueue> d_queue;
  std::vector<double> zeros(n_welford, 0.0);
  if (!snap.upload(h_snap.data(), n_state, "upload init") ||
      !sd.upload(h_sd.data(), n_state, "upload proposal_sd") ||
      !sd0.upload(h_sd.data(), n_state, "upload proposal_sd") ||
      !welford.upload(zeros.data(), n_welford, "allocate state") ||
      !temps.upload(h_temps.data(), n_chains, "allocate state") ||
      !lp.alloc(n_chains, "allocate state") ||
      !draws.alloc(n_draws, "allocate draws") ||
      !accept.alloc(n_chains, "allocate state") ||
      !acc_hist.alloc((size_t)nb * n_chains, "allocate accept history") ||
      !sync.alloc(kBatchSyncHead + 2 * (size_t)n_models, "allocate queue"))
    return 1;
  std::vector<ChainArgs> args(n_models);
  std::vector<ChainHot> hots(n_models);
  std::vector<BatchModelInfo> info(n_models);
  std::vector<BatchWork> work;
  for (uint32_t k = 0; k < n_models; ++k) {
    const gpum_batch_model& b = models[k];
    gpum_hip_model* m = b.model;
    const Tiling& t = plan[k].t;
    const uint32_t np = m->n_params, C = b.n_chains;
    ChainArgs& a = args[k];
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
    a.n_chains = C;
    a.n_params = np;
    a.init = snap.p + state_off[k];
    a.proposal_sd = sd.p + state_off[k];
    a.temperatures = temps.p + plan[k].chain_off;
    a.gamma = b.gamma;
    a.de_noise = b.de_noise;
    a.mode = o.proposal_mode;
    a.n_iter = o.n_iter;
    a.seed = b.seed;
    a.stream_id = b.stream_id;
    a.draws = draws.p + plan[k].draws_off;
    a.accept_rate = accept.p + plan[k].chain_off;
    a.last_log_post = lp.p + plan[k].chain_off;
    a.draws_ld = o.n_iter;
    a.cpb = t.cpb;
    a.wpc = t.wpc;
    a.tile_len = t.tile_len;
    a.resident = 0;
    a.lds_tile_offset = t.lds_tile_offset;
    ChainHot& h = hots[k];
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
    h.n_chains = C;
    h.n_iter = o.n_iter;
    h.cpb = t.cpb;
    h.wpc = t.wpc;
    h.tile_len = t.tile_len;
    h.unit_chunks = t.unit;
    h.stream_mode = 1;
    set_sum_fields(m, &h);
    h.lds_tile_offset = t.lds_tile_offset;
    h.state_words = 6 * np + 6;
    BatchModelInfo& in = info[k];
    in.chain_off = plan[k].chain_off;
    in.n_chains = C;
    in.n_params = np;
    in.n_groups = (C + t.cpb - 1) / t.cpb;
    in.draws_ld = o.n_iter;
    in.draws = a.draws;
    in.snapshot = snap.p + state_off[k];
    in.sd = sd.p + state_off[k];
    in.sd0 = sd0.p + state_off[k];
    in.welford = welford.p + welford_off[k];
    in.accept = accept.p + plan[k].chain_off;
    in.accept_hist = acc_hist.p + (size_t)nb * plan[k].chain_off;
    for (uint32_t c0 = 0; c0 < C; c0 += t.cpb) work.push_back(BatchWork{k, c0});
  }
  // Per-XCD queues.  The dispatcher hands workgroup w to XCD w mod 8, and
  // each of MI355X's 8 XCDs has its own L2.  A model's chain groups stream
  // the same data tiles every iteration, so all of a model's items go to
  // one XCD's queue and the tiles one of them fetches serve the others from
  // that L2.  Models are assigned largest first to the XCD with the least
  // work so far; within a queue the longest items come first in every
  // batch, so short ones fill in at the end.
  BatchQueue q;
  memset(&q, 0, sizeof q);
  {
    std::vector<uint32_t> by_size(n_models);
    for (uint32_t k = 0; k < n_models; ++k) by_size[k] = k;
    auto work_of = [&](uint32_t k) {
      return (double)models[k].model->n_obs * models[k].n_chains;
    };
    std::stable_sort(by_size.begin(), by_size.end(), [&](uint32_t x, uint32_t y) {
      return models[x].model->n_obs > models[y].model->n_obs;
    });
    std::vector<std::vector<BatchWork>> xq(kXcds);
    std::vector<double> load(kXcds, 0.0);
    for (uint32_t k : by_size) {
      const uint32_t x = (uint32_t)(std::min_element(load.begin(), load.end()) -
                                    load.begin());
      for (const BatchWork& bw : work)
        if (bw.model == k) xq[x].push_back(bw);
      load[x] += work_of(k);
    }
    work.clear();
    // Workgroups per queue: a CU's share of LDS allows one or two, and a
    // queue needs no more workgroups than it has items per batch (more
    // would only take items of later batches and wait).
    const size_t wg_lds = lds + kStreamStaticLds + kBatchArgsLds;
    const uint32_t per_cu = std::min<uint32_t>(
        2u, std::max<uint32_t>(1u, (uint32_t)(g_dev.lds_max / wg_lds)));
    const uint32_t cu_per_xcd = std::max(1, g_dev.n_cu / (int)kXcds);
    for (uint32_t x = 0; x < kXcds; ++x) {
      q.off[x] = (uint32_t)work.size();
      q.len[x] = (uint32_t)xq[x].size();
      q.slots[x] = std::min(q.len[x], cu_per_xcd * per_cu);
      work.insert(work.end(), xq[x].begin(), xq[x].end());
    }
  }
  uint32_t max_slots = 0;
  for (uint32_t x = 0; x < kXcds; ++x) max_slots = std::max(max_slots, q.slots[x]);
  const uint32_t n_wg = kXcds * max_slots;
  int clock_khz = 0;
  if (hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate,
                            g_dev.id) != hipSuccess || clock_khz <= 0)
    clock_khz = 100000;
  q.n_models = n_models;
  q.n_batches = nb;
  q.every = every;
  q.n_iter = o.n_iter;
  q.warmup_batches = o.adapt ? o.warmup_batches : 0u;
  // A dependency is another workgroup's batch of one chain group: far below
  // this bound, which only ends a run that cannot progress.
  q.wait_ticks = (uint64_t)clock_khz * 1000ull * kBatchWaitSeconds;
  if (!d_work.upload(work.data(), work.size(), "upload launch arguments") ||
      !d_info.upload(info.data(), n_models, "upload launch arguments"))
    return 1;
  q.items = d_work.p;
  q.models = d_info.p;
  q.sync = sync.p;
  if (!d_args.upload(args.data(), n_models, "upload launch arguments") ||
      !d_hots.upload(hots.data(), n_models, "upload launch arguments") ||
      !d_queue.upload(&q, 1, "upload launch arguments"))
    return 1;
  const uint32_t n_sync = kBatchSyncHead + 2 * n_models;

  hipStream_t st = g_dev.stream;
  // One compiled kernel serves the launch when every model runs the same
  // compiled program with the launch's variant.
  hipFunction_t jf = nullptr;
  {
    const std::string src0 = model_jit_source(models[0].model, force.v);
    bool same = true;
    for (uint32_t k = 0; k < n_models && same; ++k)
      same = models[k].model->jit &&
             (k == 0 || model_jit_source(models[k].model, force.v) == src0);
    if (same) {
      bool failed = false;
      jf = model_jit_kernel(models[0].model, JIT_BATCH, force.v,
                            lds + kBatchArgsLds, &failed);
      if (failed) return 1;
    }
  }
  if (!jf && !interp_has_variant(force.v)) {
    // The interpreter's batched kernels hold one or two chains per wave,
    // four for affine programs.
    set_err("chains_per_wave %u runs on compiled kernels (every model "
            "compiled, one program) or for affine programs only", cw);
    return 1;
  }
  auto enqueue = [&]() -> hipError_t {
    hipError_t e = launch_batch_sync_reset(sync.p, n_sync, st);
    if (e != hipSuccess) return e;
    if (jf) {
      const BatchQueue* qp = d_queue.p;
      const ChainArgs* ap = d_args.p;
      const ChainHot* hp = d_hots.p;
      void* kargs[] = {(void*)&qp, (void*)&ap, (void*)&hp};
      return launch_module(jf, dim3(n_wg), dim3(1024), lds, st, kargs);
    }
    return launch_chain_kernel_batch(d_queue.p, n_wg, d_args.p, d_hots.p,
                                     force.v.ns, force.v.opt, force.v.cw, lds,
                                     st);
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
    set_hip_err("batched run", e);
    return 1;
  }
  uint32_t err_word = 0;
  if (hipMemcpy(&err_word, sync.p + kXcds, sizeof err_word,
                hipMemcpyDeviceToHost) != hipSuccess) {
    set_err("read results: device error");
    return 1;
  }
  if (err_word != 0) {
    set_err("batched run: a work item waited more than %u s for its model's "
            "previous batch; the run was abandoned", kBatchWaitSeconds);
    return 1;
  }
  // Results per model.
  std::vector<float> h_draws(n_draws), h_acc((size_t)nb * n_chains);
  std::vector<double> h_lp(n_chains), h_sd_out(n_state), h_snap_out(n_state);
  bool ok = hipMemcpy(h_draws.data(), draws.p, n_draws * sizeof(float),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(h_acc.data(), acc_hist.p, h_acc.size() * sizeof(float),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(h_lp.data(), lp.p, n_chains * sizeof(double),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(h_sd_out.data(), sd.p, n_state * sizeof(double),
                      hipMemcpyDeviceToHost) == hipSuccess &&
            hipMemcpy(h_snap_out.data(), snap.p, n_state * sizeof(double),
                      hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    set_err("read results: device error");
    return 1;
  }
  for (uint32_t k = 0; k < n_models; ++k) {
    gpum_batch_model& b = models[k];
    const uint32_t C = b.n_chains, np = b.model->n_params;
    if (b.draws)
      memcpy(b.draws, h_draws.data() + plan[k].draws_off,
             (size_t)o.n_iter * C * np * sizeof(float));
    if (b.accept_hist)
      memcpy(b.accept_hist, h_acc.data() + (size_t)nb * plan[k].chain_off,
             (size_t)nb * C * sizeof(float));
    if (b.last_log_post)
      memcpy(b.last_log_post, h_lp.data() + plan[k].chain_off, C * sizeof(double));
    if (b.final_sd)
      memcpy(b.final_sd, h_sd_out.data() + state_off[k], (size_t)C * np * sizeof(double));
    if (b.final_state)
      memcpy(b.final_state, h_snap_out.data() + state_off[k],
             (size_t)C * np * sizeof(double));
  }
  opts_in->kernel_ms = ms;
  if (opts_in->struct_size >= offsetof(gpum_batch_opts, n_workgroups) + 4)
    opts_in->n_workgroups = n_wg;
  return 0;
}
