/*
 * gpum_hip.h -- C ABI of the MI355X (gfx950) sampling backend.
 *
 * The entry points mirror the reference's Rust backend interface so that a
 * `Backend::Hip` arm can call them with the arguments it already has:
 *
 *   gpum_hip_run_chains    <- gpu_metropolis_chain(...) -> ChainResult
 *                             src/rust/src/chain_kernel.rs:468-488 (result
 *                             struct :360-369), reached from
 *                             rust_gpu_metropolis src/rust/src/lib.rs:69-123
 *   gpum_hip_available     <- rust_available_backends src/rust/src/lib.rs:51-61
 *   gpum_hip_loglik_batch  <- loglik_batch src/rust/src/cpu_native.rs:339-360
 *                             via rust_loglik_batch src/rust/src/lib.rs:131-151
 *   gpum_hip_loglik_pointwise <- loglik_pointwise cpu_native.rs:371-391
 *                             via rust_loglik_pointwise lib.rs:187-207
 *   gpum_hip_grad_batch    <- grad_batch cpu_native.rs:306-329 (tape
 *                             :148-279) via rust_grad_batch lib.rs:158-179
 *
 * Conventions (as in the reference):
 *   - `code` is the DSL bytecode, (op, arg) u32 pairs (R/dsl.R:13-17);
 *     `code_len` counts u32 words.  An empty program evaluates to 0.
 *   - `data` is row-major n_obs x n_cols f64 (R/api.R:288-290).
 *   - `init`, `proposal_sd` are chain-major (n_chains x n_params);
 *     `proposal_l` is n_chains x n_params x n_params row-major lower
 *     triangular, or NULL.
 *   - proposal_mode: 0 diagonal random walk, 1 Differential Evolution against
 *     the `init` snapshot, 2 Cholesky random walk, 3 MALA (cpu_native.rs:
 *     573-724: preconditioned, truncated Langevin drift with the
 *     asymmetric-proposal correction; preconditioner proposal_l when given,
 *     else proposal_sd).  The reference's GPU backends refuse MALA
 *     (R/api.R:373-377); here it runs on the stepwise sampler with the
 *     log-posterior and its gradient in FP64 (reverse mode over the
 *     bytecode, the model's FP32 columns widened), at most 1024
 *     instructions in the program and in the prior; gpum_run_ext fused
 *     batches (swap_every) are not available with it.
 *   - draws are f32, column-major (n_iter, n_chains, n_params):
 *     draws[t + n_iter * (c + n_chains * j)].
 *   - Uniforms come from a counter-based Philox4x32-10 stream keyed on
 *     (seed, stream id) and addressed by (iteration, chain, slot); the slot
 *     layout per iteration is the reference's consumption order.  The
 *     numbers differ from the reference's triple32 stream.
 *
 * Return values: 0 on success; nonzero on failure with a message available
 * from gpum_hip_last_error() (thread-local).  Messages about allocation
 * failures contain "out of memory"; failed launches contain "kernel"; a
 * missing or unsupported GPU contains "device" -- the phrases the
 * reference's R error guard matches (R/api.R:12-34).
 *
 * Threading: calls are synchronous and may be made from any thread; the
 * library serialises its own device work per process.
 */
#ifndef GPUM_HIP_H
#define GPUM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GPUM_HIP_ABI_VERSION 1

/* chain_kernel.rs:360-369 (ChainResult).  Buffers are owned by the library
 * until gpum_hip_result_free(). */
typedef struct gpum_chain_result {
  uint32_t n_iter;
  uint32_t n_chains;
  uint32_t n_params;
  float* draws;          /* n_iter * n_chains * n_params */
  float* accept_rate;    /* n_chains */
  double* last_log_post; /* n_chains: untempered log-posterior of the final state */
} gpum_chain_result;

/* Optional extension of a run.  Zero-initialise and set struct_size;
 * fields beyond struct_size are treated as zero. */
typedef struct gpum_run_ext {
  uint32_t struct_size;
  uint32_t stream_id;     /* Philox key word 1 (model index); default 0 */
  /* Host buffers filled when non-NULL: */
  double* trace_plp;      /* n_iter * n_chains: log-posterior of each proposal */
  double* trace_prop;     /* n_iter * n_chains * n_params: each proposal */
  uint8_t* trace_accept;  /* n_iter * n_chains: 1 if accepted */
  double* final_state;    /* n_chains * n_params: state after the last iteration */
  double* init_log_post;  /* n_chains: log-posterior of `init` as evaluated */
  /* Tiling overrides (0 = autotuned). */
  int32_t chains_per_block;
  int32_t waves_per_chain;
  /* Lowering switches: -1 = default (on), 0 = off, 1 = on. */
  int32_t precompute_data;
  int32_t fuse;
  /* Out: device time of the chain kernel in milliseconds. */
  double kernel_ms;
  /* Batches inside one launch -- the parallel-tempering orchestration of
   * R/adaptive.R:272-413 fused into the kernel.  swap_every = 0: none.
   * The run is cut into max(1, n_iter / swap_every) batches (the last
   * absorbs the remainder); after each batch, chains adapt (during the first
   * adapt_batches batches: Robbins-Monro on the batch acceptance and
   * sd = scale * sqrt(diag cov), R/adaptive.R:342-356) and each ladder of
   * swap_group consecutive chains attempts the alternating adjacent swap
   * sweep (R/adaptive.R:362-391).  swap_group must divide n_chains and be at
   * most 16; swap_group <= 1 adapts without swapping. */
  uint32_t iter_offset;   /* Philox iteration counter of the first iteration */
  uint32_t swap_every;
  uint32_t swap_group;
  uint32_t adapt_batches;
  uint32_t batch0;        /* global index offset of the first batch */
  const double* init_scale;  /* in: n_chains starting RM scales, NULL = 1 */
  uint8_t* swap_hist;     /* out: n_batches * n_chains, 1/0 per attempted pair */
  float* acc_hist;        /* out: n_batches * n_chains batch acceptance */
  double* final_sd;       /* out: n_chains * n_params adapted proposal sd */
  double* final_scale;    /* out: n_chains RM scales */
  double* trace_ll;       /* out: n_iter * n_chains data term of each proposal */
  /* Chains evaluated together by one wave on every observation it reads
   * (0 = auto; 2 and 4 need a program without a stack, vm_depth 1; 4 runs
   * on the model's compiled kernel). */
  int32_t chains_per_wave;
} gpum_run_ext;

/* Static description of a lowered model. */
typedef struct gpum_model_info {
  uint32_t n_vm_instr;     /* per-observation VM program length */
  uint32_t vm_depth;       /* peak stack depth of the VM program */
  uint32_t ref_depth;      /* peak stack depth of the DSL bytecode */
  uint32_t n_uniform;      /* hoisted values + constants */
  uint32_t n_cols_total;   /* data columns streamed (incl. precomputed) */
  uint32_t n_hoist_words;  /* FP64 hoist program length (u32 words) */
  double w_exec;           /* nominal VALU slots per observation evaluation */
  double w_bytecode;       /* same weights on the program as compiled */
  /* How the kernel sums the program over observations: 0 its value, 1 the
   * square of its value before a trailing square, 2 an affine program
   * accumulated directly; n_sum_tail trailing uniform operations are
   * applied once to each data sum instead of per observation. */
  uint32_t sum_mode;
  uint32_t n_sum_tail;
} gpum_model_info;

typedef struct gpum_hip_model gpum_hip_model;

/* 1 if a gfx950 device is visible and usable, else 0. */
int gpum_hip_available(void);

/* Message of the last failed call on this thread ("" if none). */
const char* gpum_hip_last_error(void);

/* One batch of whole-chain Metropolis iterations; the HIP replacement of
 * gpu_metropolis_chain (chain_kernel.rs:468-532). */
int gpum_hip_run_chains(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        const uint32_t* prior_code, size_t prior_len,
                        const double* prior_consts, size_t n_prior_consts,
                        uint32_t n_params, const double* data, uint32_t n_cols,
                        uint32_t n_obs, const double* init, uint32_t n_chains,
                        const double* proposal_sd, const double* temperatures,
                        uint32_t n_iter, uint32_t seed, uint32_t proposal_mode,
                        double gamma, double de_noise,
                        const double* proposal_l, gpum_chain_result* out);

/* Same, with the extension struct (may be NULL). */
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
                           gpum_run_ext* ext, gpum_chain_result* out);

/* Free the buffers of a result filled by a run call. */
void gpum_hip_result_free(gpum_chain_result* r);

/* Log-likelihood summed over observations at each of n_points parameter
 * vectors (point-major), cpu_native.rs:339-360.  out: n_points doubles. */
int gpum_hip_loglik_batch(const uint32_t* code, size_t code_len,
                          const double* consts, size_t n_consts,
                          uint32_t n_params, const double* data,
                          uint32_t n_cols, uint32_t n_obs,
                          const double* points, uint32_t n_points,
                          double* out);

/* Log-likelihood of each observation at each of n_points parameter vectors,
 * evaluated in FP64 on the device; the HIP replacement of loglik_pointwise
 * (cpu_native.rs:371-391, FFI rust_loglik_pointwise lib.rs:187-207), which
 * backs WAIC and PSIS-LOO.  out: n_points * n_obs doubles, out[p * n_obs + i]
 * for observation i at point p.  As the reference, n_params = 0 means no
 * points and an empty program evaluates to 0. */
int gpum_hip_loglik_pointwise(const uint32_t* code, size_t code_len,
                              const double* consts, size_t n_consts,
                              uint32_t n_params, const double* data,
                              uint32_t n_cols, uint32_t n_obs,
                              const double* points, uint32_t n_points,
                              double* out);

/* Gradient of the log-likelihood (data term, summed over the observations)
 * at each of n_points parameter vectors, by reverse-mode differentiation of
 * the bytecode in FP64 on the device; the HIP replacement of grad_batch
 * (cpu_native.rs:148-279, 306-329, FFI rust_grad_batch lib.rs:158-179),
 * which backs the Cramer-Rao diagnostic.  out: n_points * n_params doubles,
 * point-major.  At most 16 parameters and 1024 instructions. */
int gpum_hip_grad_batch(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        uint32_t n_params, const double* data, uint32_t n_cols,
                        uint32_t n_obs, const double* points, uint32_t n_points,
                        double* out);

/* gpum_hip_grad_batch without device allocation.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (code under HIP graph
 * capture, serving loops, harnesses that forbid it).  `d_data` (FP64,
 * column-major: column c at d_data[c * n_obs]), `d_points` (point-major) and
 * `d_out` (n_points * n_params, point-major) are DEVICE buffers the caller
 * owns; the program's tape, its constants and the partial sums live in an
 * 8 MiB array of the library, one per GPU, held until the process ends.
 * Limits beyond gpum_hip_grad_batch's: at most 4096 constants, and
 * ceil(n_obs / 8192) * (n_params + 1) * 8 bytes of partial sums per point
 * must fit in the array (about 2.4e8 observations at 16 parameters); larger
 * calls return an error.  Points are processed in chunks that fit.  Results
 * equal gpum_hip_grad_batch's bit for bit (the interpreted kernels; this
 * entry never compiles a per-program kernel, which would allocate).  Calls
 * are serialised by the library's lock; every call reuses the array, so
 * calls from two processes sharing one GPU each have their own copy, and
 * the array keeps the last call's contents.  Returns after the results are
 * in d_out. */
int gpum_hip_grad_batch_noalloc(const uint32_t* code, size_t code_len,
                                const double* consts, size_t n_consts,
                                uint32_t n_params, const double* d_data,
                                uint32_t n_cols, uint32_t n_obs,
                                const double* d_points, uint32_t n_points,
                                double* d_out);

/* Device time (ms) of the kernels of this thread's last
 * gpum_hip_loglik_pointwise or gpum_hip_grad_batch call (HIP events on the
 * library's stream; copies excluded).  No counterpart in the reference. */
double gpum_hip_last_eval_ms(void);

/* 1 if that call ran kernels compiled for its program (hiprtc, cached per
 * process by program), 0 if it ran the interpreter.  No counterpart in the
 * reference. */
int gpum_hip_last_eval_compiled(void);

/* Let the FP64 evaluation -- gpum_hip_loglik_pointwise,
 * gpum_hip_grad_batch and the gradient of MALA (proposal_mode 3) -- compile
 * kernels per program (1, the default; it falls back to the interpreter
 * where the process cannot) or always interpret (0).  Both give the same
 * values, operation for operation.  Process-wide.  No counterpart in the
 * reference. */
void gpum_hip_set_eval_jit(int32_t enable);

/* A model with its lowered program and data resident on the device, for
 * repeated batches without re-uploading the data (the reference uploads
 * every call, chain_kernel.rs:393-413).  precompute_data / fuse: -1
 * default, 0 off, 1 on.  Returns NULL on failure. */
gpum_hip_model* gpum_hip_model_create(
    const uint32_t* code, size_t code_len, const double* consts,
    size_t n_consts, const uint32_t* prior_code, size_t prior_len,
    const double* prior_consts, size_t n_prior_consts, uint32_t n_params,
    const double* data, uint32_t n_cols, uint32_t n_obs,
    int32_t precompute_data, int32_t fuse);

int gpum_hip_model_info(const gpum_hip_model* m, gpum_model_info* info);

/* What the kernels would run for a log-likelihood program, without a device:
 * the lowering of gpum_hip_model_create on the host only (validation,
 * hoisting, precomputed columns, fusion, summation fold, common
 * subexpressions).  No counterpart in the reference; for inspecting a model
 * before it is uploaded.  fuse / precompute_data as gpum_hip_model_create. */
int gpum_hip_lower_info(const uint32_t* code, size_t code_len,
                        const double* consts, size_t n_consts,
                        uint32_t n_params, uint32_t n_cols,
                        int32_t precompute_data, int32_t fuse,
                        gpum_model_info* info);

/* Compile, with hiprtc and without a device, the kernels
 * gpum_hip_model_set_jit would compile for this program and prior with
 * chains_per_wave chains per wave (0 or 1: one).  kinds: bit k selects
 * kernel kind k (0 chain streaming, 1 chain resident, 2 points streaming,
 * 3 points resident, 4 batched, 5 FP64 pointwise, 6 FP64 gradient).
 * compile_ms (may be NULL): total compile
 * time.  No counterpart in the reference; it checks that a model's compiled
 * kernels build on a host without a GPU. */
int gpum_hip_jit_check(const uint32_t* code, size_t code_len,
                       const double* consts, size_t n_consts,
                       const uint32_t* prior_code, size_t prior_len,
                       const double* prior_consts, size_t n_prior_consts,
                       uint32_t n_params, uint32_t n_cols,
                       int32_t chains_per_wave, uint32_t kinds,
                       double* compile_ms);

/* One batch on a resident model; arguments as gpum_hip_run_chains_ex. */
int gpum_hip_model_run(gpum_hip_model* m, const double* init,
                       uint32_t n_chains, const double* proposal_sd,
                       const double* temperatures, uint32_t n_iter,
                       uint32_t seed, uint32_t proposal_mode, double gamma,
                       double de_noise, const double* proposal_l,
                       gpum_run_ext* ext, gpum_chain_result* out);

/* ---- Adaptive warmup on the device -------------------------------------
 * Replaces the host orchestration .am_orchestrate_warmup (R/adaptive.R:
 * 101-254, called from R/api.R:531-536) for random-walk chains, with its
 * batch plan (R/adaptive.R:112-118: 10 to 40 batches of about 50
 * iterations), per-chain and pooled Welford moments, Robbins-Monro scalar
 * per chain toward the target acceptance of R/adaptive.R:69-72, restart
 * of every accumulator at mid-warmup, proposal L_c = scale_c * the
 * regularised Cholesky factor of the pooled covariance (R/adaptive.R:548-
 * 558) from the second batch on, and optional early freeze (auto_stop,
 * R/adaptive.R:224-239).  Every batch and every between-batch step runs on
 * the device (with use_graph, as one HIP graph); batch b draws with seed
 * + b - 1.  init and proposal_sd: n_chains * n_params, chain-major. */
typedef struct gpum_warmup_opts {
  uint32_t struct_size;
  uint32_t warmup;            /* iterations, > 0 */
  uint32_t seed;
  int32_t auto_stop;
  int32_t use_graph;
  /* host outputs, each may be NULL: */
  double* final_init;         /* n_chains * n_params: the last draws */
  double* final_sd;           /* n_chains * n_params */
  double* final_scale;        /* n_chains: Robbins-Monro scalars */
  double* final_l;            /* n_chains * n_params^2, row-major per chain;
                                 meaningful when has_l */
  float* accept_hist;         /* min(40, warmup) * n_chains, batch-major;
                                 rows past n_batches are unspecified */
  uint32_t n_batches_planned; /* out */
  uint32_t n_batches;         /* out: batches run (fewer after a freeze) */
  uint32_t warmup_used;       /* out: iterations run */
  uint32_t has_l;             /* out: 1 when final_l holds a factor */
  double kernel_ms;           /* out: device time of the warmup */
} gpum_warmup_opts;

int gpum_hip_model_warmup(gpum_hip_model* m, const double* init,
                          uint32_t n_chains, const double* proposal_sd,
                          gpum_warmup_opts* opts);

/* Loglik batch on a resident model. */
int gpum_hip_model_loglik(gpum_hip_model* m, const double* points,
                          uint32_t n_points, double* out);

/* Run this model on kernels compiled for its program (enable = 1) or on
 * the interpreter (0, the default).  The lowered program becomes
 * straight-line code calling the interpreter's own per-instruction
 * arithmetic, compiled at run time for gfx950 (hiprtc) and cached per
 * process by program; results equal the interpreter's.  No counterpart on
 * the reference's GPU path; its CPU path compiles each model the same way
 * (src/rust/src/jit.rs:441-631).  compile_ms (may be NULL): compile time of
 * the first kernel, 0 when cached. */
int gpum_hip_model_set_jit(gpum_hip_model* m, int32_t enable,
                           double* compile_ms);

void gpum_hip_model_free(gpum_hip_model* m);

/* ---- Surrogate delayed acceptance -------------------------------------
 * No counterpart in the reference: these entry points add a learned first
 * stage to the random-walk sampler.  The surrogate is an MLP
 * n_in(<=16, zero-padded) - 256 - 256 - 1 with SiLU activations, trained on
 * the GPU (bf16 matrix-core products, FP32 accumulation, Adam on FP32
 * weights) to predict a model's summed log-likelihood from a parameter
 * vector.  Inputs are whitened, x = W (theta - mean); outputs are
 * loglik = target_mean + target_sd * network(x). */
typedef struct gpum_hip_surrogate gpum_hip_surrogate;

/* Number of FP32 weights of the network (the flat vector's length):
 * W1 16x256 | b1 256 | W2 256x256 | b2 256 | w3 256 | b3 1, row-major
 * (input index first). */
uint32_t gpum_hip_surrogate_n_weights(void);

/* A network for n_in <= 16 inputs, weights drawn from `seed`. */
gpum_hip_surrogate* gpum_hip_surrogate_create(uint32_t n_in, uint32_t seed);
void gpum_hip_surrogate_free(gpum_hip_surrogate* s);

/* mean: n_in; whiten: n_in x n_in row-major. */
int gpum_hip_surrogate_set_scaling(gpum_hip_surrogate* s, const double* mean,
                                   const double* whiten, double target_mean,
                                   double target_sd);
int gpum_hip_surrogate_get_weights(gpum_hip_surrogate* s, float* out);
int gpum_hip_surrogate_set_weights(gpum_hip_surrogate* s, const float* w);

/* Gradient of the mean squared error of the standardised prediction,
 * mean_i (network(x_i) - (loglik_i - target_mean) / target_sd)^2, over n
 * points (n <= 65536), with respect to the flat weights.  grad_out:
 * gpum_hip_surrogate_n_weights() floats; loss_out: the loss. */
int gpum_hip_surrogate_gradient(gpum_hip_surrogate* s, const double* theta,
                                const double* loglik, uint32_t n,
                                float* grad_out, double* loss_out);

typedef struct gpum_train_opts {
  uint32_t struct_size;
  uint32_t steps;     /* Adam steps */
  uint32_t batch;     /* rows per step (multiple of 64, <= 65536) */
  uint32_t seed;      /* minibatch sampling stream */
  double lr;          /* initial learning rate */
  double lr_final;    /* cosine decay target */
  double beta1;       /* Adam moment decays and epsilon */
  double beta2;
  double eps;
} gpum_train_opts;

/* Train on n (theta, loglik) pairs; loss_trace (steps doubles, may be NULL)
 * receives each step's minibatch loss; kernel_ms (may be NULL) the device
 * time of the whole training. */
int gpum_hip_surrogate_train(gpum_hip_surrogate* s, const double* theta,
                             const double* loglik, uint32_t n,
                             const gpum_train_opts* opts, double* loss_trace,
                             double* kernel_ms);

/* Predicted log-likelihood at n points. */
int gpum_hip_surrogate_predict(gpum_hip_surrogate* s, const double* theta,
                               uint32_t n, double* out);

/* Extension of a stepwise run.  Zero-initialise and set struct_size. */
typedef struct gpum_da_ext {
  uint32_t struct_size;
  uint32_t stream_id;
  uint32_t iter_offset;
  uint32_t work_items;      /* exact-stage work items per launch (0 auto) */
  uint32_t work_groups;     /* exact-stage persistent grid size (0 auto) */
  double* final_state;      /* out: n_chains * n_params */
  double* stage1_rate;      /* out: n_chains fraction reaching the exact stage */
  double* trace_s;          /* out: n_iter * n_chains surrogate log-posterior */
  double* trace_plp;        /* out: n_iter * n_chains exact (NaN: not run) */
  uint8_t* trace_stage;     /* out: 0 stage-1 reject, 1 stage-2 reject, 2 accept */
  double kernel_ms;         /* out: device time of the iterations */
  double exact_evals;       /* out: proposals evaluated exactly */
  double* trace_prop;       /* out: n_iter * n_chains * n_params proposals */
  double* trace_ll;         /* out: n_iter * n_chains data term (NaN: not run) */
  uint32_t chains_per_wave; /* exact-stage slots per wave (0 auto, 1, 2 or
                               4; 2 and 4 need vm_depth 1, 4 a compiled
                               kernel) */
  double gamma;             /* proposal_mode 1: difference scale (> 0) */
  double de_noise;          /* proposal_mode 1: jitter scale on proposal_sd */
} gpum_da_ext;

/* Random-walk Metropolis (proposal_mode 0 or 2), synchronous Differential
 * Evolution (proposal_mode 1, ext->gamma / ext->de_noise, at least 3
 * chains, no surrogate: path B of cpu_native.rs:400-506, every chain
 * proposing against the previous generation's whole population; the launch
 * boundary is the generation barrier) or MALA (proposal_mode 3, no
 * surrogate; see proposal_mode above) run one iteration per launch
 * sequence.  With a surrogate (s non-NULL, n_in == n_params), each
 * proposal first faces the surrogate's Metropolis test and only survivors
 * are evaluated exactly; the exact test divides out the surrogate ratio,
 * so the target is unchanged (two-stage delayed acceptance).  With
 * s == NULL every proposal is evaluated exactly, with the uniforms of
 * gpum_hip_model_run.  Result layout as gpum_hip_run_chains. */
int gpum_hip_model_run_da(gpum_hip_model* m, gpum_hip_surrogate* s,
                          const double* init, uint32_t n_chains,
                          const double* proposal_sd,
                          const double* temperatures, uint32_t n_iter,
                          uint32_t seed, uint32_t proposal_mode,
                          const double* proposal_l, gpum_da_ext* ext,
                          gpum_chain_result* out);

/* Synchronous Differential Evolution MCMC (path B), the HIP replacement of
 * run_de_sync (cpu_native.rs:400-506, FFI rust_gpu_metropolis_de_sync
 * lib.rs:215-260), which the reference runs on its CPU path only
 * (R/api.R:464-472).  Each generation every chain proposes
 * x + g (x_a - x_b) + de_noise * proposal_sd * z against the previous
 * generation, a != c and b not in {c, a}, g = 1 with probability 0.1 else
 * gamma, and accepts when log u < plp - cur.  Uniforms: the Philox stream
 * of gpum_hip_run_chains, 2 n_params + 4 slots per iteration in the
 * reference's order.  n_chains >= 3.  Result layout as gpum_hip_run_chains. */
int gpum_hip_run_de_sync(const uint32_t* code, size_t code_len,
                         const double* consts, size_t n_consts,
                         const uint32_t* prior_code, size_t prior_len,
                         const double* prior_consts, size_t n_prior_consts,
                         uint32_t n_params, const double* data, uint32_t n_cols,
                         uint32_t n_obs, const double* init, uint32_t n_chains,
                         const double* proposal_sd, uint32_t n_iter,
                         uint32_t seed, double gamma, double de_noise,
                         gpum_chain_result* out);

/* ---- Many models in one run -------------------------------------------
 * No single-call counterpart in the reference, whose host runs one model
 * per rust_gpu_metropolis call (R/api.R:247-599).  Each model keeps its own
 * program, data, chains, seed and Philox stream id.  One persistent launch
 * runs every batch of iterations of every model: each chain group of each
 * model is a work item per batch, an item starts once its model has ended
 * the previous batch, and the workgroup that ends a model's batch runs the
 * between-batch step of the reference's orchestrator for it
 * (R/adaptive.R:427-524 for DE path A: the population snapshot is
 * refreshed from the last f32 draws, and during the warmup batches each
 * chain's sd becomes sqrt(diag cov) of its draws so far).  With use_graph
 * the launch (and the reset of its queue counters) is one HIP graph.
 * Batch b (1-based) draws its uniforms with seed + b - 1, as the reference
 * reseeds each batch (R/adaptive.R:471).  A run in which an item waits
 * more than 30 s for its model's previous batch is abandoned with an
 * error. */
typedef struct gpum_batch_model {
  gpum_hip_model* model;
  uint32_t n_chains;
  const double* init;         /* n_chains * n_params */
  const double* proposal_sd;  /* n_chains * n_params */
  uint32_t seed;
  uint32_t stream_id;         /* Philox key word 1 */
  double gamma;               /* DE difference scale (proposal_mode 1) */
  double de_noise;
  /* Host outputs filled when non-NULL: */
  float* draws;               /* n_iter * n_chains * n_params, all iterations,
                                 draws[t + n_iter * (c + n_chains * j)] */
  float* accept_hist;         /* n_batches * n_chains */
  double* final_sd;           /* n_chains * n_params */
  double* last_log_post;      /* n_chains */
  double* final_state;        /* n_chains * n_params: the last snapshot */
} gpum_batch_model;

typedef struct gpum_batch_opts {
  uint32_t struct_size;
  uint32_t proposal_mode;     /* 0 random walk, 1 DE against the snapshot */
  uint32_t n_iter;            /* all iterations, warmup included */
  uint32_t batch_len;         /* iterations per batch (de_every); the last
                                 batch absorbs the remainder */
  uint32_t warmup_batches;    /* leading batches that adapt */
  int32_t adapt;
  int32_t chains_per_block;   /* 0 = 16 * chains_per_wave;
                                 chains_per_block / chains_per_wave *
                                 waves_per_chain = 16 */
  int32_t waves_per_chain;
  int32_t use_graph;          /* launch through a HIP graph */
  double kernel_ms;           /* out: device time of the run */
  uint32_t n_workgroups;      /* out: workgroups of the persistent launch */
  int32_t chains_per_wave;    /* 0 or 1: one chain per wave; 2: two (every
                                 model of vm_depth 1); 4: four (also every
                                 model compiled, one program) */
} gpum_batch_opts;

int gpum_hip_batch_run(gpum_batch_model* models, uint32_t n_models,
                       gpum_batch_opts* opts);

#ifdef __cplusplus
}
#endif
#endif /* GPUM_HIP_H */
