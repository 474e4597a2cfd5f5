// da.cpp -- host side of the surrogate network and of the stepwise sampler
// (plain or two-stage delayed-acceptance Metropolis).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/gpum_hip.h"
#include "da.h"
#include "jit.h"
#include "mlp.h"
#include "philox.h"
#include "runtime.h"

using namespace gpum;
using namespace gpum::rt;

struct gpum_hip_surrogate {
  uint32_t n_in = 0;
  DevBuf<float> P, G, Ma, Va;
  DevBuf<uint16_t> W1b, W2b;
  std::vector<double> mean, whiten;
  double t_mean = 0.0, t_sd = 1.0;
  DevBuf<double> d_mean, d_whiten;
  uint32_t adam_t = 0;
  // workspace for `cap` rows (a multiple of 64)
  uint32_t cap = 0;
  DevBuf<uint16_t> Xb, H1b, D2b, D1b;
  DevBuf<float> A1, A2, H2, s, tb, part, cs_b1, loss;
};

namespace {

bool sur_workspace(gpum_hip_surrogate* S, uint32_t rows) {
  rows = round_up(std::max(rows, 64u), 64u);
  if (rows <= S->cap) return true;
  const size_t r = rows, h = kMlpHid;
  if (!S->Xb.alloc(r * kMlpIn, "allocate surrogate workspace") ||
      !S->H1b.alloc(r * h, "allocate surrogate workspace") ||
      !S->D2b.alloc(r * h, "allocate surrogate workspace") ||
      !S->D1b.alloc(r * h, "allocate surrogate workspace") ||
      !S->A1.alloc(r * h, "allocate surrogate workspace") ||
      !S->A2.alloc(r * h, "allocate surrogate workspace") ||
      !S->H2.alloc(r * h, "allocate surrogate workspace") ||
      !S->s.alloc(r, "allocate surrogate workspace") ||
      !S->tb.alloc(r, "allocate surrogate workspace") ||
      !S->part.alloc((r / 64) * kPartStride, "allocate surrogate workspace") ||
      !S->cs_b1.alloc((r / 16) * h, "allocate surrogate workspace") ||
      !S->loss.alloc(1, "allocate surrogate workspace"))
    return false;
  // Rows past a call's count are read as zero inputs.
  HIP_TRY("clear surrogate workspace",
          hipMemsetAsync(S->Xb.p, 0, r * kMlpIn * sizeof(uint16_t), g_dev.stream));
  S->cap = rows;
  return true;
}

// Forward pass over `rows` rows of S->Xb (a multiple of 16); with `train`
// the pre-activations are kept for the backward pass.
bool sur_forward(gpum_hip_surrogate* S, uint32_t rows, bool train,
                 hipStream_t st) {
  GemmArgs g;
  memset(&g, 0, sizeof g);
  g.A = S->Xb.p, g.sam = kMlpIn, g.sak = 1;
  g.B = S->W1b.p, g.sbk = kMlpHid, g.sbn = 1;
  g.M = (int)rows, g.N = kMlpHid, g.K = kMlpIn;
  g.bias = S->P.p + kOffB1;
  g.out = train ? S->A1.p : nullptr;
  g.act_b = S->H1b.p;
  g.ld = kMlpHid;
  HIP_TRY("surrogate layer 1", launch_gemm_bf16(g, st));
  memset(&g, 0, sizeof g);
  g.A = S->H1b.p, g.sam = kMlpHid, g.sak = 1;
  g.B = S->W2b.p, g.sbk = kMlpHid, g.sbn = 1;
  g.M = (int)rows, g.N = kMlpHid, g.K = kMlpHid;
  g.bias = S->P.p + kOffB2;
  g.out = train ? S->A2.p : nullptr;
  g.act = S->H2.p;
  g.ld = kMlpHid;
  HIP_TRY("surrogate layer 2", launch_gemm_bf16(g, st));
  HIP_TRY("surrogate head", launch_mlp_head(S->H2.p, S->P.p, S->s.p, (int)rows, st));
  return true;
}

// Backward pass of the loss over the first `valid` of `rows` rows (a
// multiple of 64) into S->G; the summed squared error lands in S->loss
// (and, in a training step with `sched`, in losses[k] with the step's rate
// set and the counters advanced).
bool sur_backward(gpum_hip_surrogate* S, uint32_t rows, uint32_t valid,
                  hipStream_t st, MlpSched* sched = nullptr,
                  float* losses = nullptr, const MlpTrainConst* tc = nullptr) {
  HIP_TRY("surrogate head backward",
          launch_mlp_head_bwd(S->s.p, S->tb.p, S->H2.p, S->A2.p, S->P.p,
                              S->D2b.p, S->part.p, (int)rows, (int)valid, st));
  GemmArgs g;
  // dW2 = H1^T D2
  memset(&g, 0, sizeof g);
  g.A = S->H1b.p, g.sam = 1, g.sak = kMlpHid;
  g.B = S->D2b.p, g.sbk = kMlpHid, g.sbn = 1;
  g.M = kMlpHid, g.N = kMlpHid, g.K = (int)rows;
  g.out = S->G.p + kOffW2;
  g.ld = kMlpHid;
  HIP_TRY("surrogate dW2", launch_gemm_bf16(g, st));
  // D1 = (D2 W2^T) * silu'(A1), column sums for b1
  memset(&g, 0, sizeof g);
  g.A = S->D2b.p, g.sam = kMlpHid, g.sak = 1;
  g.B = S->W2b.p, g.sbk = 1, g.sbn = kMlpHid;
  g.M = (int)rows, g.N = kMlpHid, g.K = kMlpHid;
  g.dact = S->A1.p;
  g.out_b = S->D1b.p;
  g.colsum = S->cs_b1.p;
  g.ld = kMlpHid;
  HIP_TRY("surrogate delta1", launch_gemm_bf16(g, st));
  // dW1 = X^T D1
  memset(&g, 0, sizeof g);
  g.A = S->Xb.p, g.sam = 1, g.sak = kMlpIn;
  g.B = S->D1b.p, g.sbk = kMlpHid, g.sbn = 1;
  g.M = kMlpIn, g.N = kMlpHid, g.K = (int)rows;
  g.out = S->G.p + kOffW1;
  g.ld = kMlpHid;
  HIP_TRY("surrogate dW1", launch_gemm_bf16(g, st));
  HIP_TRY("surrogate gradient",
          launch_mlp_grad_finish(S->part.p, S->cs_b1.p, (int)rows, S->G.p,
                                 S->loss.p, st, sched, losses, tc));
  return true;
}

// Network inputs of n host points into S->Xb.
bool sur_features(gpum_hip_surrogate* S, const double* theta, uint32_t n,
                  DevBuf<double>* tmp, hipStream_t st) {
  if (!tmp->upload(theta, (size_t)n * S->n_in, "upload points")) return false;
  HIP_TRY("surrogate features",
          launch_da_features(tmp->p, n, S->n_in, S->d_mean.p, S->d_whiten.p,
                             S->Xb.p, st));
  return true;
}

bool sur_check(gpum_hip_surrogate* S) {
  if (!S) {
    set_err("null surrogate");
    return false;
  }
  if (!device_init()) return false;
  hipSetDevice(g_dev.id);
  return true;
}

}  // namespace

extern "C" {

uint32_t gpum_hip_surrogate_n_weights(void) { return (uint32_t)kMlpParams; }

gpum_hip_surrogate* gpum_hip_surrogate_create(uint32_t n_in, uint32_t seed) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return nullptr;
  hipSetDevice(g_dev.id);
  if (n_in == 0 || n_in > (uint32_t)kMlpIn) {
    set_err("surrogate inputs must be in [1, %d] (got %u)", kMlpIn, n_in);
    return nullptr;
  }
  gpum_hip_surrogate* S = new gpum_hip_surrogate();
  S->n_in = n_in;
  // LeCun-uniform weights (variance 1 / fan_in), zero biases, from Philox
  // blocks (i, 0, 1, TAG_MLP).
  std::vector<float> w(kMlpParams, 0.0f);
  auto fill = [&](int off, int count, int fan_in) {
    const double a = sqrt(3.0 / (double)fan_in);
    for (int i = 0; i < count; ++i) {
      U4 h = philox4x32_10(U4{(uint32_t)(off + i), 0u, 1u, TAG_MLP}, seed, 0u);
      w[off + i] = (float)(a * (2.0 * u32_to_unit(h.x) - 1.0));
    }
  };
  fill(kOffW1, kMlpIn * kMlpHid, (int)n_in);
  // Inputs past n_in are always zero; their rows of W1 stay zero.
  for (int k = (int)n_in; k < kMlpIn; ++k)
    for (int j = 0; j < kMlpHid; ++j) w[kOffW1 + k * kMlpHid + j] = 0.0f;
  fill(kOffW2, kMlpHid * kMlpHid, kMlpHid);
  fill(kOffW3, kMlpHid, kMlpHid);
  S->mean.assign(n_in, 0.0);
  S->whiten.assign((size_t)n_in * n_in, 0.0);
  for (uint32_t k = 0; k < n_in; ++k) S->whiten[k * n_in + k] = 1.0;
  std::vector<float> zero(kMlpParams, 0.0f);
  bool ok = S->P.upload(w.data(), w.size(), "upload surrogate weights") &&
            S->G.upload(zero.data(), zero.size(), "allocate surrogate") &&
            S->Ma.upload(zero.data(), zero.size(), "allocate surrogate") &&
            S->Va.upload(zero.data(), zero.size(), "allocate surrogate") &&
            S->W1b.alloc((size_t)kMlpIn * kMlpHid, "allocate surrogate") &&
            S->W2b.alloc((size_t)kMlpHid * kMlpHid, "allocate surrogate") &&
            S->d_mean.upload(S->mean.data(), n_in, "allocate surrogate") &&
            S->d_whiten.upload(S->whiten.data(), S->whiten.size(), "allocate surrogate");
  if (ok) {
    hipError_t e = launch_mlp_mirror(S->P.p, S->W1b.p, S->W2b.p, g_dev.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(g_dev.stream);
    if (e != hipSuccess) {
      set_hip_err("surrogate init", e);
      ok = false;
    }
  }
  if (!ok) {
    delete S;
    return nullptr;
  }
  return S;
}

void gpum_hip_surrogate_free(gpum_hip_surrogate* S) {
  std::lock_guard<std::mutex> lk(g_mu);
  delete S;
}

int gpum_hip_surrogate_set_scaling(gpum_hip_surrogate* S, const double* mean,
                                   const double* whiten, double target_mean,
                                   double target_sd) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  if (!mean || !whiten) {
    set_err("mean and whiten must not be NULL");
    return 1;
  }
  if (!(target_sd > 0.0) || !isfinite(target_sd) || !isfinite(target_mean)) {
    set_err("target_sd must be positive and finite");
    return 1;
  }
  const uint32_t n = S->n_in;
  S->mean.assign(mean, mean + n);
  S->whiten.assign(whiten, whiten + (size_t)n * n);
  S->t_mean = target_mean;
  S->t_sd = target_sd;
  if (!S->d_mean.upload(S->mean.data(), n, "upload scaling") ||
      !S->d_whiten.upload(S->whiten.data(), S->whiten.size(), "upload scaling"))
    return 1;
  return 0;
}

int gpum_hip_surrogate_get_weights(gpum_hip_surrogate* S, float* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  if (hipMemcpy(out, S->P.p, kMlpParams * sizeof(float), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    set_err("read surrogate weights: device error");
    return 1;
  }
  return 0;
}

int gpum_hip_surrogate_set_weights(gpum_hip_surrogate* S, const float* w) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  hipError_t e = hipMemcpy(S->P.p, w, kMlpParams * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_mlp_mirror(S->P.p, S->W1b.p, S->W2b.p, g_dev.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_dev.stream);
  if (e != hipSuccess) {
    set_hip_err("write surrogate weights", e);
    return 1;
  }
  S->adam_t = 0;
  e = hipMemsetAsync(S->Ma.p, 0, kMlpParams * sizeof(float), g_dev.stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(S->Va.p, 0, kMlpParams * sizeof(float), g_dev.stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g_dev.stream);
  if (e != hipSuccess) {
    set_hip_err("reset optimiser", e);
    return 1;
  }
  return 0;
}

int gpum_hip_surrogate_gradient(gpum_hip_surrogate* S, const double* theta,
                                const double* loglik, uint32_t n,
                                float* grad_out, double* loss_out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  if (n == 0 || n > 65536 || !theta || !loglik || !grad_out) {
    set_err("gradient needs 1..65536 points and non-NULL buffers");
    return 1;
  }
  const uint32_t rows = round_up(n, 64);
  if (!sur_workspace(S, rows)) return 1;
  hipStream_t st = g_dev.stream;
  DevBuf<double> tmp;
  if (!sur_features(S, theta, n, &tmp, st)) return 1;
  std::vector<float> t(rows, 0.0f);
  for (uint32_t i = 0; i < n; ++i) t[i] = (float)((loglik[i] - S->t_mean) / S->t_sd);
  if (hipMemcpyAsync(S->tb.p, t.data(), rows * sizeof(float), hipMemcpyHostToDevice,
                     st) != hipSuccess) {
    set_err("upload targets: device error");
    return 1;
  }
  if (!sur_forward(S, rows, true, st) || !sur_backward(S, rows, n, st)) return 1;
  float se = 0.0f;
  hipError_t e = hipMemcpyAsync(grad_out, S->G.p, kMlpParams * sizeof(float),
                                hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(&se, S->loss.p, sizeof(float), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) {
    set_hip_err("surrogate gradient", e);
    return 1;
  }
  if (loss_out) *loss_out = (double)se / (double)n;
  return 0;
}

int gpum_hip_surrogate_train(gpum_hip_surrogate* S, const double* theta,
                             const double* loglik, uint32_t n,
                             const gpum_train_opts* opts, double* loss_trace,
                             double* kernel_ms) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  gpum_train_opts o;
  memset(&o, 0, sizeof o);
  if (opts) memcpy(&o, opts, std::min<size_t>(opts->struct_size, sizeof o));
  if (o.steps == 0) o.steps = 2000;
  if (o.batch == 0) o.batch = 1024;
  if (!(o.lr > 0.0)) o.lr = 2e-3;
  if (!(o.lr_final >= 0.0)) o.lr_final = 0.0;
  if (o.beta1 == 0.0) o.beta1 = 0.9;
  if (o.beta2 == 0.0) o.beta2 = 0.999;
  if (o.eps == 0.0) o.eps = 1e-8;
  if (n == 0 || !theta || !loglik) {
    set_err("training needs at least one (theta, loglik) pair");
    return 1;
  }
  if (o.batch % 64 || o.batch > 65536) {
    set_err("batch must be a multiple of 64 and at most 65536");
    return 1;
  }
  if (!sur_workspace(S, o.batch)) return 1;
  hipStream_t st = g_dev.stream;
  // The training set as network inputs (bf16 rows) and standardised targets.
  DevBuf<double> tmp;
  DevBuf<uint16_t> Xtr;
  DevBuf<float> Ttr, losses;
  if (!tmp.upload(theta, (size_t)n * S->n_in, "upload training set") ||
      !Xtr.alloc((size_t)n * kMlpIn, "allocate training set") ||
      !losses.alloc(o.steps, "allocate loss trace"))
    return 1;
  std::vector<float> t(n);
  for (uint32_t i = 0; i < n; ++i) t[i] = (float)((loglik[i] - S->t_mean) / S->t_sd);
  if (!Ttr.upload(t.data(), n, "upload training set")) return 1;
  hipError_t fe = launch_da_features(tmp.p, n, S->n_in, S->d_mean.p,
                                     S->d_whiten.p, Xtr.p, st);
  if (fe != hipSuccess) {
    set_hip_err("surrogate features", fe);
    return 1;
  }
  // The step counters and the rate live on the device, so every step is
  // the same sequence of launches: chunks of kGraphSteps steps are captured
  // as one HIP graph and replayed (a step is ten small launches; launched
  // one by one, their latency, not the matrix cores, bounds training).
  DevBuf<MlpSched> sched;
  {
    MlpSched s0;
    memset(&s0, 0, sizeof s0);
    s0.t = S->adam_t;
    if (!sched.upload(&s0, 1, "allocate training schedule")) return 1;
  }
  MlpTrainConst tc;
  memset(&tc, 0, sizeof tc);
  tc.steps = o.steps;
  tc.beta1 = (float)o.beta1;
  tc.beta2 = (float)o.beta2;
  tc.eps = (float)o.eps;
  tc.lr = o.lr;
  tc.lr_final = o.lr_final;
  const uint32_t B = o.batch;
  const uint32_t valid = std::min(B, n);
  auto step = [&]() -> bool {
    return launch_mlp_gather(Xtr.p, Ttr.p, n, o.seed, 0, S->Xb.p, S->tb.p,
                             (int)B, (int)valid, st, sched.p) == hipSuccess &&
           sur_forward(S, B, true, st) &&
           sur_backward(S, B, valid, st, sched.p, losses.p, &tc) &&
           launch_adam_sched(S->P.p, S->G.p, S->Ma.p, S->Va.p, S->W1b.p,
                             S->W2b.p, sched.p, (float)o.beta1, (float)o.beta2,
                             (float)o.eps, st) == hipSuccess;
  };
  constexpr uint32_t kGraphSteps = 50;
  const uint32_t chunk = std::min(kGraphSteps, o.steps);
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool ok = true;
  hipError_t e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
  if (e == hipSuccess) {
    for (uint32_t k = 0; k < chunk && ok; ++k) ok = step();
    e = hipStreamEndCapture(st, &graph);
  }
  if (ok && e == hipSuccess)
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  if (ok && e == hipSuccess) e = hipEventRecord(e0, st);
  for (uint32_t k = 0; ok && e == hipSuccess && k + chunk <= o.steps; k += chunk)
    e = hipGraphLaunch(exec, st);
  for (uint32_t k = o.steps / chunk * chunk; ok && e == hipSuccess && k < o.steps;
       ++k)
    ok = step();
  if (ok && e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0.0f;
  if (ok && e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (exec) hipGraphExecDestroy(exec);
  if (graph) hipGraphDestroy(graph);
  if (!ok || e != hipSuccess) {
    if (ok) set_hip_err("surrogate training", e);
    return 1;
  }
  S->adam_t += o.steps;
  if (kernel_ms) *kernel_ms = ms;
  if (loss_trace) {
    std::vector<float> l(o.steps);
    if (hipMemcpy(l.data(), losses.p, o.steps * sizeof(float), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      set_err("read loss trace: device error");
      return 1;
    }
    for (uint32_t k = 0; k < o.steps; ++k) loss_trace[k] = (double)l[k] / (double)valid;
  }
  return 0;
}

int gpum_hip_surrogate_predict(gpum_hip_surrogate* S, const double* theta,
                               uint32_t n, double* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!sur_check(S)) return 1;
  if (n == 0) return 0;
  if (!theta || !out) {
    set_err("theta and out must not be NULL");
    return 1;
  }
  const uint32_t chunk = 16384;
  if (!sur_workspace(S, std::min(round_up(n, 64), chunk))) return 1;
  hipStream_t st = g_dev.stream;
  DevBuf<double> tmp;
  std::vector<float> s(chunk);
  for (uint32_t i0 = 0; i0 < n; i0 += chunk) {
    const uint32_t k = std::min(chunk, n - i0);
    if (!sur_features(S, theta + (size_t)i0 * S->n_in, k, &tmp, st)) return 1;
    if (!sur_forward(S, round_up(k, 16), false, st)) return 1;
    hipError_t e = hipMemcpyAsync(s.data(), S->s.p, k * sizeof(float),
                                  hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) {
      set_hip_err("surrogate predict", e);
      return 1;
    }
    for (uint32_t i = 0; i < k; ++i) out[i0 + i] = S->t_mean + S->t_sd * (double)s[i];
  }
  return 0;
}

// Langevin steps (proposal_mode 3, cpu_native.rs:573-724), with the
// library lock held: the log-posterior and its gradient come from the FP64
// tape of the reference programs over the model's columns; every
// iteration draws the proposals, evaluates value and gradient at all of
// them (observations split over every CU) and decides.
static int run_mala_locked(gpum_hip_model* m, const double* init, uint32_t C,
                           const double* proposal_sd,
                           const double* temperatures, uint32_t n_iter,
                           uint32_t seed, const double* proposal_l,
                           const gpum_da_ext& X, gpum_da_ext* ext_in,
                           gpum_chain_result* out) {
  const uint32_t np = m->n_params;
  if (m->code_ref.size() / 2 > kMaxTapeOps) {
    set_err("Langevin proposals differentiate programs of at most %u "
            "instructions", kMaxTapeOps);
    return 1;
  }
  if (m->prior_ref.size() / 2 > kMaxTapeOps) {
    set_err("Langevin proposals differentiate priors of at most %u "
            "instructions", kMaxTapeOps);
    return 1;
  }
  if (!ensure_eval64(m)) return 1;
  const uint32_t n_ops = (uint32_t)(m->code_ref.size() / 2);
  const uint32_t ns = (m->n_obs + kGradSliceObs - 1) / kGradSliceObs;
  const size_t ncp = (size_t)C * np;
  const size_t n_draws = (size_t)n_iter * ncp;
  DevBuf<double> theta, prop, cur, acc, passes, temps, sd, L, gx, gy, mean_x,
      ll, part, vpr, gpr, ppart;
  DevBuf<float> draws, draws_it;
  DevBuf<double> tr_plp, tr_prop, tr_ll;
  DevBuf<uint8_t> tr_stage;
  DevBuf<DaArgs> d_args;
  bool ok = theta.upload(init, ncp, "upload init") &&
            prop.alloc(ncp, "allocate state") && cur.alloc(C, "allocate state") &&
            acc.alloc(C, "allocate state") && passes.alloc(C, "allocate state") &&
            gx.alloc(ncp, "allocate state") && gy.alloc(ncp, "allocate state") &&
            mean_x.alloc(ncp, "allocate state") && ll.alloc(C, "allocate state") &&
            part.alloc(std::max<size_t>(1, (size_t)C * ns * (np + 1)),
                       "allocate partial sums") &&
            vpr.alloc(C, "allocate state") && gpr.alloc(ncp, "allocate state") &&
            ppart.alloc((size_t)C * (np + 1), "allocate partial sums") &&
            temps.upload(temperatures, C, "upload temperatures") &&
            sd.upload(proposal_sd, ncp, "upload proposal_sd") &&
            draws.alloc(n_draws, "allocate draws") &&
            draws_it.alloc(n_draws, "allocate draws");
  if (ok && proposal_l) ok = L.upload(proposal_l, ncp * np, "upload proposal_l");
  if (ok && X.trace_plp) ok = tr_plp.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && X.trace_stage) ok = tr_stage.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && X.trace_prop) ok = tr_prop.alloc(n_draws, "allocate trace");
  if (ok && X.trace_ll) ok = tr_ll.alloc((size_t)n_iter * C, "allocate trace");
  if (!ok) return 1;
  hipStream_t st = g_dev.stream;
  if (hipMemsetAsync(acc.p, 0, C * sizeof(double), st) != hipSuccess ||
      hipMemsetAsync(passes.p, 0, C * sizeof(double), st) != hipSuccess) {
    set_err("clear counters: device error");
    return 1;
  }
  DaArgs a;
  memset(&a, 0, sizeof a);
  a.n_chains = C;
  a.n_params = np;
  a.n_iter = n_iter;
  a.seed = seed;
  a.stream_id = X.stream_id;
  a.iter_offset = X.iter_offset;
  a.mode = 3;
  a.temperatures = temps.p;
  a.proposal_sd = sd.p;
  a.proposal_l = proposal_l ? L.p : nullptr;
  a.prior_consts = m->prior_consts.p;
  a.theta = theta.p;
  a.prop = prop.p;
  a.cur = cur.p;
  a.accepts = acc.p;
  a.passes = passes.p;
  a.gx = gx.p;
  a.gy = gy.p;
  a.mean_x = mean_x.p;
  a.ll_prop = ll.p;
  a.v_prior = vpr.p;
  a.g_prior = gpr.p;
  a.draws = draws.p;
  a.draws_it = draws_it.p;
  a.trace_plp = X.trace_plp ? tr_plp.p : nullptr;
  a.trace_stage = X.trace_stage ? tr_stage.p : nullptr;
  a.trace_prop = X.trace_prop ? tr_prop.p : nullptr;
  a.trace_ll = X.trace_ll ? tr_ll.p : nullptr;
  if (!d_args.upload(&a, 1, "upload launch arguments")) return 1;

  // Value and gradient of the data term at C points (the states or the
  // proposals) into ll / g.
  Eval64Args ea{};
  ea.prog = m->tape_ll.p;
  ea.n_ops = n_ops;
  ea.n_params = np;
  ea.consts = m->consts64.p;
  ea.data = m->data64.p;
  ea.n_obs = m->n_obs;
  ea.n_points = C;
  ea.out = part.p;
  ea.n_slices = ns;
  ea.depth = (uint32_t)m->L.ref_depth;
  ea.n_slots = m->slots_ll;
  const uint32_t grid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(grad64_items(ns, C), (uint64_t)g_dev.n_cu * 8u));
  // The gradient compiled for the program where the process can, as the
  // batch evaluation's (same values either way).
  const hipFunction_t gjf = eval64_compiled_kernel(
      m->code_ref.data(), m->code_ref.size(), m->consts_ref.data(), JIT_GRAD64);
  // The prior: a program without data, one "observation" per point.
  const uint32_t n_prior_ops = (uint32_t)(m->prior_ref.size() / 2);
  Eval64Args pe{};
  pe.prog = m->tape_prior.p;
  pe.n_ops = n_prior_ops;
  pe.n_params = np;
  pe.consts = m->prior_consts.p;
  pe.n_obs = 1;
  pe.n_points = C;
  pe.out = ppart.p;
  pe.n_slices = 1;
  pe.depth = m->prior_depth;
  pe.n_slots = m->slots_prior;
  const hipFunction_t pjf = eval64_compiled_kernel(
      m->prior_ref.data(), m->prior_ref.size(), m->prior_consts_h.data(),
      JIT_GRAD64);
  const uint32_t pgrid = (uint32_t)std::max<uint64_t>(
      1, std::min<uint64_t>(grad64_items(1, C), (uint64_t)g_dev.n_cu * 8u));
  auto prior_term = [&](const double* points) -> hipError_t {
    if (n_prior_ops == 0) {
      hipError_t r = hipMemsetAsync(vpr.p, 0, C * sizeof(double), st);
      return r == hipSuccess ? hipMemsetAsync(gpr.p, 0, ncp * sizeof(double), st)
                             : r;
    }
    Eval64Args e = pe;
    e.points = points;
    hipError_t r;
    if (pjf) {
      void* kargs[] = {(void*)&e};
      r = launch_module(pjf, dim3(pgrid), dim3(256), 0, st, kargs);
    } else {
      r = launch_grad64(e, pgrid, st);
    }
    return r == hipSuccess
               ? launch_grad64_finish(ppart.p, C, np, 1, gpr.p, vpr.p, st)
               : r;
  };
  auto data_term = [&](const double* points, double* g) -> hipError_t {
    if (n_ops == 0) {
      hipError_t r = hipMemsetAsync(ll.p, 0, C * sizeof(double), st);
      return r == hipSuccess ? hipMemsetAsync(g, 0, ncp * sizeof(double), st) : r;
    }
    Eval64Args e = ea;
    e.points = points;
    hipError_t r;
    if (gjf) {
      void* kargs[] = {(void*)&e};
      r = launch_module(gjf, dim3(grid), dim3(256), 0, st, kargs);
    } else {
      r = launch_grad64(e, grid, st);
    }
    return r == hipSuccess ? launch_grad64_finish(part.p, C, np, ns, g, ll.p, st)
                           : r;
  };
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  chk(data_term(theta.p, gy.p));
  chk(prior_term(theta.p));
  chk(launch_mala_init(d_args.p, C, st));
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, st);
  for (uint32_t it = 0; it < n_iter && e == hipSuccess; ++it) {
    chk(launch_mala_propose(d_args.p, C, it, st));
    chk(data_term(prop.p, gy.p));
    chk(prior_term(prop.p));
    chk(launch_mala_accept(d_args.p, C, it, st));
  }
  if (e == hipSuccess) e = launch_da_draws(draws_it.p, draws.p, n_iter, C * np, st);
  hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0.0f;
  if (e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) {
    set_hip_err("Langevin sampler", e);
    return 1;
  }
  std::vector<double> h_acc(C);
  out->n_iter = n_iter;
  out->n_chains = C;
  out->n_params = np;
  out->draws = (float*)malloc(std::max<size_t>(1, n_draws) * sizeof(float));
  out->accept_rate = (float*)malloc(C * sizeof(float));
  out->last_log_post = (double*)malloc(C * sizeof(double));
  if (!out->draws || !out->accept_rate || !out->last_log_post) {
    gpum_hip_result_free(out);
    set_err("host out of memory for results");
    return 1;
  }
  auto d2h = [](void* dst, const void* src, size_t bytes) {
    return hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
  };
  ok = d2h(out->draws, draws.p, n_draws * sizeof(float)) &&
       d2h(out->last_log_post, cur.p, C * sizeof(double)) &&
       d2h(h_acc.data(), acc.p, C * sizeof(double));
  if (ok && X.final_state) ok = d2h(X.final_state, theta.p, ncp * sizeof(double));
  if (ok && X.trace_plp)
    ok = d2h(X.trace_plp, tr_plp.p, (size_t)n_iter * C * sizeof(double));
  if (ok && X.trace_stage) ok = d2h(X.trace_stage, tr_stage.p, (size_t)n_iter * C);
  if (ok && X.trace_prop) ok = d2h(X.trace_prop, tr_prop.p, n_draws * sizeof(double));
  if (ok && X.trace_ll)
    ok = d2h(X.trace_ll, tr_ll.p, (size_t)n_iter * C * sizeof(double));
  if (!ok) {
    gpum_hip_result_free(out);
    set_err("read results: device error");
    return 1;
  }
  for (uint32_t c = 0; c < C; ++c) {
    out->accept_rate[c] = n_iter == 0 ? 0.0f : (float)h_acc[c] / (float)n_iter;
    if (X.stage1_rate) X.stage1_rate[c] = n_iter == 0 ? 0.0 : 1.0;
  }
  if (ext_in) {
    if (ext_in->struct_size >= offsetof(gpum_da_ext, kernel_ms) + sizeof(double))
      ext_in->kernel_ms = ms;
    if (ext_in->struct_size >= offsetof(gpum_da_ext, exact_evals) + sizeof(double))
      ext_in->exact_evals = (double)n_iter * C;
  }
  return 0;
}

// The stepwise sampler with the library lock held.
static int run_da_locked(gpum_hip_model* m, gpum_hip_surrogate* S,
                         const double* init, uint32_t n_chains,
                         const double* proposal_sd,
                         const double* temperatures, uint32_t n_iter,
                         uint32_t seed, uint32_t proposal_mode,
                         const double* proposal_l, gpum_da_ext* ext_in,
                         gpum_chain_result* out) {
  if (!m || !out) {
    set_err("null model or result");
    return 1;
  }
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  gpum_da_ext X;
  memset(&X, 0, sizeof X);
  if (ext_in) memcpy(&X, ext_in, std::min<size_t>(ext_in->struct_size, sizeof X));
  const uint32_t np = m->n_params;
  const uint32_t C = n_chains;
  if (C == 0 || !init || !temperatures) {
    set_err("n_chains must be positive; init and temperatures non-NULL");
    return 1;
  }
  if (proposal_mode > 3) {
    set_err("the stepwise sampler runs proposal_mode 0, 1, 2 or 3 (got %u)",
            proposal_mode);
    return 1;
  }
  if ((proposal_mode != 2 && !proposal_sd) || (proposal_mode == 2 && !proposal_l)) {
    set_err("proposal_mode %u needs %s", proposal_mode,
            proposal_mode == 2 ? "proposal_l" : "proposal_sd");
    return 1;
  }
  if (proposal_mode == 3) {
    // Langevin: preconditioner proposal_l when given, else proposal_sd.
    if (S) {
      set_err("Langevin proposals run without a surrogate");
      return 1;
    }
    for (uint32_t c = 0; c < C; ++c) {
      if (!(temperatures[c] > 0.0) || !isfinite(temperatures[c])) {
        set_err("temperatures must be positive and finite");
        return 1;
      }
    }
    return run_mala_locked(m, init, C, proposal_sd, temperatures, n_iter, seed,
                           proposal_l, X, ext_in, out);
  }
  if (proposal_mode == 1) {
    // Differential Evolution, path B: the pair of other chains must exist.
    if (C < 3) {
      set_err("synchronous Differential Evolution needs at least 3 chains");
      return 1;
    }
    if (!(X.gamma > 0.0) || !isfinite(X.gamma) || !(X.de_noise >= 0.0) ||
        !isfinite(X.de_noise)) {
      set_err("gamma must be positive and finite, de_noise non-negative and "
              "finite");
      return 1;
    }
    if (S) {
      set_err("Differential Evolution runs without a surrogate");
      return 1;
    }
  }
  for (uint32_t c = 0; c < C; ++c) {
    if (!(temperatures[c] > 0.0) || !isfinite(temperatures[c])) {
      set_err("temperatures must be positive and finite");
      return 1;
    }
  }
  if (S && S->n_in != np) {
    set_err("surrogate has %u inputs but the model %u parameters", S->n_in, np);
    return 1;
  }
  Tiling t;
  if (!pick_tiling(m, C, nullptr, &t, true, nullptr, (int)X.chains_per_wave))
    return 1;
  if (!ensure_data(m, t.n_pad)) return 1;
  const uint32_t n_tiles = t.n_pad / t.tile_len;
  // Exact-stage work items per launch and grid: by default one item per
  // workgroup the CUs hold at once (160 KB of LDS and 32 waves a CU), so an
  // item's start (queue, uniforms, first tile) is paid once per resident
  // workgroup.  On 1024 DE chains over 1e6 rows (one 148 KB workgroup a
  // CU) 256 items took 0.090 ms a generation, 512 items 0.095 to 0.103.
  const size_t wg_lds = t.lds_bytes + kPointsExtraLds + 64 +
                        (t.stream_mode ? kStreamStaticLds : 0);
  const uint32_t per_cu = std::max<uint32_t>(
      1u, std::min<uint32_t>((uint32_t)(g_dev.lds_max / wg_lds),
                             32u / std::max(1u, tiling_waves(t))));
  const uint32_t resident = per_cu * (uint32_t)g_dev.n_cu;
  const uint32_t items = X.work_items ? X.work_items : resident;
  const uint32_t max_slices = std::max(1u, std::min(n_tiles, items));
  const uint32_t groups_cap = (C + t.cpb - 1) / t.cpb;
  const uint32_t n_wg =
      X.work_groups
          ? X.work_groups
          : std::max(1u, (uint32_t)std::min<uint64_t>(
                             (uint64_t)groups_cap * n_tiles, resident));

  const size_t ncp = (size_t)C * np;
  const size_t n_draws = (size_t)n_iter * ncp;
  DevBuf<double> theta, prop, cur, cur_s, pprior, prop_s, acc, passes, partial,
      temps, sd, L;
  DevBuf<int32_t> pos;
  DevBuf<uint32_t> list, count, queue, n_sl;
  DevBuf<float> draws, draws_it;
  DevBuf<double> tr_s, tr_plp, tr_prop, tr_ll;
  DevBuf<uint8_t> tr_stage;
  DevBuf<DaArgs> d_args;
  bool ok = theta.upload(init, ncp, "upload init") &&
            prop.alloc(ncp, "allocate state") && cur.alloc(C, "allocate state") &&
            cur_s.alloc(C, "allocate state") && pprior.alloc(C, "allocate state") &&
            prop_s.alloc(C, "allocate state") && acc.alloc(C, "allocate state") &&
            passes.alloc(C, "allocate state") &&
            partial.alloc((size_t)max_slices * C, "allocate partials") &&
            queue.alloc(2, "allocate state") && n_sl.alloc(1, "allocate state") &&
            temps.upload(temperatures, C, "upload temperatures") &&
            pos.alloc(C, "allocate state") && list.alloc(C, "allocate state") &&
            count.alloc(1, "allocate state") &&
            draws.alloc(n_draws, "allocate draws") &&
            draws_it.alloc(n_draws, "allocate draws");
  if (ok && proposal_mode != 2) ok = sd.upload(proposal_sd, ncp, "upload proposal_sd");
  if (ok && proposal_mode == 2)
    ok = L.upload(proposal_l, ncp * np, "upload proposal_l");
  if (ok && X.trace_s) ok = tr_s.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && X.trace_plp) ok = tr_plp.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && X.trace_stage) ok = tr_stage.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && X.trace_prop) ok = tr_prop.alloc(n_draws, "allocate trace");
  if (ok && X.trace_ll) ok = tr_ll.alloc((size_t)n_iter * C, "allocate trace");
  if (ok && S) ok = sur_workspace(S, round_up(C, 64));
  if (!ok) return 1;
  if (hipMemsetAsync(queue.p, 0, 2 * sizeof(uint32_t), g_dev.stream) != hipSuccess) {
    set_err("clear work queue: device error");
    return 1;
  }
  if (X.trace_s && hipMemsetAsync(tr_s.p, 0xFF, (size_t)n_iter * C * sizeof(double),
                                  g_dev.stream) != hipSuccess) {
    set_err("clear trace: device error");
    return 1;
  }

  DaArgs a;
  memset(&a, 0, sizeof a);
  a.n_chains = C;
  a.n_params = np;
  a.n_iter = n_iter;
  a.seed = seed;
  a.stream_id = X.stream_id;
  a.iter_offset = X.iter_offset;
  a.mode = proposal_mode;
  a.gamma = X.gamma;
  a.de_noise = X.de_noise;
  a.temperatures = temps.p;
  a.proposal_sd = sd.p;
  a.proposal_l = L.p;
  a.prior = m->prior.p;
  a.n_prior = (uint32_t)m->prior_code.size();
  a.prior_consts = m->prior_consts.p;
  a.theta = theta.p;
  a.prop = prop.p;
  a.cur = cur.p;
  a.cur_s = cur_s.p;
  a.prop_prior = pprior.p;
  a.prop_s = prop_s.p;
  a.accepts = acc.p;
  a.passes = passes.p;
  a.pos = pos.p;
  a.list = list.p;
  a.count = count.p;
  a.partial = partial.p;
  a.n_slices = n_sl.p;
  a.use_surrogate = S ? 1u : 0u;
  if (S) {
    a.n_in = S->n_in;
    a.norm_mean = S->d_mean.p;
    a.norm_w = S->d_whiten.p;
    a.t_mean = S->t_mean;
    a.t_sd = S->t_sd;
    a.Xb = S->Xb.p;
    a.s_raw = S->s.p;
  }
  a.draws = draws.p;
  a.draws_it = draws_it.p;
  a.trace_s = X.trace_s ? tr_s.p : nullptr;
  a.trace_plp = X.trace_plp ? tr_plp.p : nullptr;
  a.trace_stage = X.trace_stage ? tr_stage.p : nullptr;
  a.trace_prop = X.trace_prop ? tr_prop.p : nullptr;
  a.trace_ll = X.trace_ll ? tr_ll.p : nullptr;
  if (!d_args.upload(&a, 1, "upload launch arguments")) return 1;

  ChainHot h;
  memset(&h, 0, sizeof h);
  h.vm = reinterpret_cast<const VmIns4*>(m->vm.p);
  h.u_init = m->u_init.p;
  h.data = m->data.p;
  h.n_vm = m->L.empty ? 0 : (uint32_t)m->L.vm.size();
  h.n_uniform_pad = m->n_uniform_pad;
  h.n_cols = m->n_cols_total;
  h.n_obs = m->n_obs;
  h.n_pad = t.n_pad;
  h.ld = m->data_ld;
  h.n_chains = C;
  h.cpb = t.cpb;
  h.wpc = 1;
  h.tile_len = t.tile_len;
  h.unit_chunks = t.unit;
  h.stream_mode = t.stream_mode;
  set_sum_fields(m, &h);
  h.lds_tile_offset = t.lds_tile_offset;
  PointsArgs pa;
  memset(&pa, 0, sizeof pa);
  pa.n_points = C;
  pa.n_params = np;
  pa.hoist = m->hoist.p;
  pa.n_hoist = (uint32_t)m->L.hoist.size();
  pa.consts = m->consts.p;
  pa.target_items = items;
  pa.queue = queue.p;
  pa.n_slices_out = n_sl.p;
  pa.partial = partial.p;

  hipStream_t st = g_dev.stream;
  const uint32_t rows16 = round_up(C, 16);
  hipError_t e = hipSuccess;
  auto chk = [&](hipError_t r) {
    if (e == hipSuccess) e = r;
  };
  // Initial states: exact log-posterior and surrogate value.
  chk(launch_da_init(d_args.p, C, st));
  if (S && e == hipSuccess && !sur_forward(S, rows16, false, st)) return 1;
  // Device copies of the points-kernel arguments: the initial states
  // (every slot), then the compacted proposals of each iteration.
  PointsArgs pinit = pa;
  pinit.points = theta.p;
  pa.points = prop.p;
  pa.index = list.p;
  pa.count = count.p;
  DevBuf<PointsArgs> d_pinit, d_pa;
  if (!d_pinit.upload(&pinit, 1, "upload launch arguments") ||
      !d_pa.upload(&pa, 1, "upload launch arguments"))
    return 1;
  bool jit_failed = false;
  hipFunction_t jf = model_jit_kernel(
      m, t.stream_mode ? JIT_POINTS_STREAM : JIT_POINTS_RESIDENT, t.v,
      t.lds_bytes + kPointsExtraLds, &jit_failed);
  if (jit_failed) return 1;
  if (!jf && !interp_has_variant(t.v)) {
    set_err("%d chains per wave need a compiled kernel, which this launch "
            "(%zu B of LDS) cannot use", t.v.cw, (size_t)t.lds_bytes);
    return 1;
  }
  auto points = [&](const PointsArgs* p) -> hipError_t {
    if (!jf)
      return launch_points_kernel(p, n_wg, h, t.v.ns, t.v.opt, t.v.cw,
                                  t.lds_bytes, st);
    void* kargs[] = {(void*)&p, (void*)&h};
    return launch_module(jf, dim3(n_wg), dim3(tiling_waves(t) * 64),
                         t.lds_bytes, st, kargs);
  };
  chk(points(d_pinit.p));
  chk(launch_da_init_finish(d_args.p, C, st));
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  hipEventRecord(e0, st);
  for (uint32_t it = 0; it < n_iter && e == hipSuccess; ++it) {
    chk(launch_da_propose(d_args.p, C, it, st));
    if (S && e == hipSuccess && !sur_forward(S, rows16, false, st)) {
      hipEventDestroy(e0);
      hipEventDestroy(e1);
      return 1;
    }
    if (S) chk(launch_da_stage1(d_args.p, it, st));
    chk(points(d_pa.p));
    chk(launch_da_stage2(d_args.p, C, it, st));
  }
  if (e == hipSuccess) e = launch_da_draws(draws_it.p, draws.p, n_iter, C * np, st);
  hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  float ms = 0.0f;
  if (e == hipSuccess) hipEventElapsedTime(&ms, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) {
    set_hip_err("stepwise sampler", e);
    return 1;
  }
  std::vector<double> h_acc(C), h_pass(C);
  out->n_iter = n_iter;
  out->n_chains = C;
  out->n_params = np;
  out->draws = (float*)malloc(std::max<size_t>(1, n_draws) * sizeof(float));
  out->accept_rate = (float*)malloc(C * sizeof(float));
  out->last_log_post = (double*)malloc(C * sizeof(double));
  if (!out->draws || !out->accept_rate || !out->last_log_post) {
    gpum_hip_result_free(out);
    set_err("host out of memory for results");
    return 1;
  }
  ok = hipMemcpy(out->draws, draws.p, n_draws * sizeof(float), hipMemcpyDeviceToHost) ==
           hipSuccess &&
       hipMemcpy(out->last_log_post, cur.p, C * sizeof(double), hipMemcpyDeviceToHost) ==
           hipSuccess &&
       hipMemcpy(h_acc.data(), acc.p, C * sizeof(double), hipMemcpyDeviceToHost) ==
           hipSuccess &&
       hipMemcpy(h_pass.data(), passes.p, C * sizeof(double), hipMemcpyDeviceToHost) ==
           hipSuccess;
  if (ok && X.final_state)
    ok = hipMemcpy(X.final_state, theta.p, ncp * sizeof(double), hipMemcpyDeviceToHost) ==
         hipSuccess;
  if (ok && X.trace_s)
    ok = hipMemcpy(X.trace_s, tr_s.p, (size_t)n_iter * C * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && X.trace_plp)
    ok = hipMemcpy(X.trace_plp, tr_plp.p, (size_t)n_iter * C * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && X.trace_stage)
    ok = hipMemcpy(X.trace_stage, tr_stage.p, (size_t)n_iter * C,
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && X.trace_prop)
    ok = hipMemcpy(X.trace_prop, tr_prop.p, n_draws * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (ok && X.trace_ll)
    ok = hipMemcpy(X.trace_ll, tr_ll.p, (size_t)n_iter * C * sizeof(double),
                   hipMemcpyDeviceToHost) == hipSuccess;
  if (!ok) {
    gpum_hip_result_free(out);
    set_err("read results: device error");
    return 1;
  }
  double evals = 0.0;
  for (uint32_t c = 0; c < C; ++c) {
    out->accept_rate[c] = n_iter == 0 ? 0.0f : (float)h_acc[c] / (float)n_iter;
    if (X.stage1_rate) X.stage1_rate[c] = n_iter == 0 ? 0.0 : h_pass[c] / n_iter;
    evals += h_pass[c];
  }
  if (ext_in) {
    gpum_da_ext* o = ext_in;
    if (o->struct_size >= offsetof(gpum_da_ext, kernel_ms) + sizeof(double))
      o->kernel_ms = ms;
    if (o->struct_size >= offsetof(gpum_da_ext, exact_evals) + sizeof(double))
      o->exact_evals = evals;
  }
  return 0;
}

}  // extern "C"

namespace gpum {
namespace rt {

int run_langevin(gpum_hip_model* m, const double* init, uint32_t n_chains,
                 const double* proposal_sd, const double* temperatures,
                 uint32_t n_iter, uint32_t seed, const double* proposal_l,
                 gpum_run_ext* ext, gpum_chain_result* out) {
  gpum_da_ext X;
  memset(&X, 0, sizeof X);
  X.struct_size = sizeof X;
  std::vector<uint8_t> stage;
  auto has = [&](size_t end) { return ext && ext->struct_size >= end; };
  if (ext) {
    if (has(offsetof(gpum_run_ext, swap_every) + 4) && ext->swap_every > 0) {
      set_err("proposal_mode 3 runs without fused batches (swap_every = 0)");
      return 1;
    }
    X.stream_id = ext->stream_id;
    X.final_state = ext->final_state;
    X.trace_plp = ext->trace_plp;
    X.trace_prop = ext->trace_prop;
    if (has(offsetof(gpum_run_ext, iter_offset) + 4)) X.iter_offset = ext->iter_offset;
    if (has(offsetof(gpum_run_ext, trace_ll) + sizeof(double*)))
      X.trace_ll = ext->trace_ll;
    if (ext->trace_accept) {
      stage.assign((size_t)n_iter * n_chains, 0);
      X.trace_stage = stage.data();
    }
  }
  const int rc = run_da_locked(m, nullptr, init, n_chains, proposal_sd,
                               temperatures, n_iter, seed, 3, proposal_l, &X,
                               out);
  if (rc == 0 && ext) {
    if (has(offsetof(gpum_run_ext, kernel_ms) + sizeof(double)))
      ext->kernel_ms = X.kernel_ms;
    for (size_t k = 0; k < stage.size(); ++k)
      ext->trace_accept[k] = stage[k] == 2 ? 1 : 0;
  }
  return rc;
}

}  // namespace rt
}  // namespace gpum

extern "C" {

int gpum_hip_model_run_da(gpum_hip_model* m, gpum_hip_surrogate* S,
                          const double* init, uint32_t n_chains,
                          const double* proposal_sd,
                          const double* temperatures, uint32_t n_iter,
                          uint32_t seed, uint32_t proposal_mode,
                          const double* proposal_l, gpum_da_ext* ext_in,
                          gpum_chain_result* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  return run_da_locked(m, S, init, n_chains, proposal_sd, temperatures, n_iter,
                       seed, proposal_mode, proposal_l, ext_in, out);
}

int gpum_hip_run_de_sync(const uint32_t* code, size_t code_len,
                         const double* consts, size_t n_consts,
                         const uint32_t* prior_code, size_t prior_len,
                         const double* prior_consts, size_t n_prior_consts,
                         uint32_t n_params, const double* data, uint32_t n_cols,
                         uint32_t n_obs, const double* init, uint32_t n_chains,
                         const double* proposal_sd, uint32_t n_iter,
                         uint32_t seed, double gamma, double de_noise,
                         gpum_chain_result* out) {
  std::lock_guard<std::mutex> lk(g_mu);
  if (!device_init()) return 1;
  hipSetDevice(g_dev.id);
  gpum_hip_model m;
  if (!model_build(&m, code, code_len, consts, n_consts, prior_code, prior_len,
                   prior_consts, n_prior_consts, n_params, data, n_cols, n_obs,
                   -1, -1))
    return 1;
  // Path B has no temperatures: log u < plp - cur (cpu_native.rs:482).
  std::vector<double> temps(n_chains, 1.0);
  gpum_da_ext X;
  memset(&X, 0, sizeof X);
  X.struct_size = sizeof X;
  X.gamma = gamma;
  X.de_noise = de_noise;
  return run_da_locked(&m, nullptr, init, n_chains, proposal_sd, temps.data(),
                       n_iter, seed, 1, nullptr, &X, out);
}

}  // extern "C"
