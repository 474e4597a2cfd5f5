/*
 * oracle.h -- CPU restatement of the reference sampler's f64 path.
 * TEST INFRASTRUCTURE ONLY (see oracle.c).
 */
#ifndef GPUM_ORACLE_H
#define GPUM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_PARAMS 64

enum { ORC_STREAM_T32 = 0, ORC_STREAM_PHILOX = 1 };

typedef double (*orc_loglik_fn)(const double* params, const double* data,
                                uint32_t n_obs, uint32_t n_cols);

typedef struct {
  const uint32_t* code;
  uint32_t code_len; /* number of u32 words (2 per instruction) */
  const double* consts;
  const uint32_t* prior_code;
  uint32_t prior_len;
  const double* prior_consts;
  uint32_t n_params;
  const double* data; /* row-major n_obs x n_cols */
  uint32_t n_cols;
  uint32_t n_obs;
  const double* init; /* chain-major n_chains x n_params */
  uint32_t n_chains;
  const double* proposal_sd;  /* chain-major */
  const double* temperatures; /* n_chains */
  uint32_t n_iter;
  uint32_t seed;
  uint32_t proposal_mode;
  double gamma;
  double de_noise;
  const double* proposal_l; /* n_chains x np x np row-major, or NULL */
  uint32_t stream;          /* ORC_STREAM_T32 or ORC_STREAM_PHILOX */
  uint32_t stream_id;       /* Philox key word 1 (model index) */
  uint32_t n_threads;
  orc_loglik_fn loglik_fn;  /* compiled data term, or NULL to interpret */
  uint32_t iter_offset;     /* Philox iteration counter of iteration 0 */
} orc_run_args;

typedef struct {
  float* draws;           /* n_iter x n_chains x n_params, column-major */
  float* accept_rate;     /* n_chains */
  double* last_log_post;  /* n_chains */
  double* final_state;    /* n_chains x n_params, or NULL */
  double* trace_plp;      /* n_iter x n_chains, or NULL */
  double* trace_prop;     /* n_iter x n_chains x n_params, or NULL */
  uint8_t* trace_accept;  /* n_iter x n_chains, or NULL */
} orc_run_out;

uint32_t orc_triple32(uint32_t h);
double orc_uniform_t32(uint32_t seedmix, uint32_t ctr);
uint32_t orc_seedmix(uint32_t seed, uint32_t chain);
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]);
double orc_uniform_philox(uint32_t seed, uint32_t stream, uint32_t tag,
                          uint32_t t, uint32_t chain, uint32_t slot);
double orc_lgamma(double x);
double orc_digamma(double x);
double orc_vm_eval(const uint32_t* code, uint32_t n_words,
                   const double* consts, const double* params,
                   const double* row);
double orc_loglik(const uint32_t* code, uint32_t n_words, const double* consts,
                  const double* params, const double* data, uint32_t n_cols,
                  uint32_t n_obs);
double orc_vm_eval_grad(const uint32_t* code, uint32_t n_words,
                        const double* consts, const double* params,
                        const double* row, double* vals, double* adj,
                        uint32_t* lhs, uint32_t* rhs, double* grad);
void orc_loglik_batch(const uint32_t* code, uint32_t n_words,
                      const double* consts, uint32_t n_params,
                      const double* data, uint32_t n_cols, uint32_t n_obs,
                      const double* points, uint32_t n_points, double* out);
void orc_loglik_pointwise(const uint32_t* code, uint32_t n_words,
                          const double* consts, uint32_t n_params,
                          const double* data, uint32_t n_cols, uint32_t n_obs,
                          const double* points, uint32_t n_points,
                          double* out);
void orc_grad_batch(const uint32_t* code, uint32_t n_words,
                    const double* consts, uint32_t n_params,
                    const double* data, uint32_t n_cols, uint32_t n_obs,
                    const double* points, uint32_t n_points, double* out);
int orc_run(const orc_run_args* a, orc_run_out* o);
int orc_run_de_sync(const orc_run_args* a, orc_run_out* o);

#ifdef __cplusplus
}
#endif
#endif
