/*
 * oracle.c -- CPU restatement of the reference sampler's f64 path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle for the HIP
 * backend: only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
 * leg load it, and only as the checker (or the timed CPU baseline).  The
 * product path never calls into it.
 *
 * Each function cites the reference file:line it restates.  All citations
 * are relative to the reference repository root (pcbrom/gpumetropolis).
 *
 * What is restated:
 *   - triple32 hash and its uniform          src/rust/src/cpu_native.rs:13-28
 *   - Lanczos log-gamma, digamma             src/rust/src/cpu_native.rs:34-77
 *   - f64 stack-machine interpreter          src/rust/src/cpu_native.rs:80-107
 *   - sequential observation sum             src/rust/src/cpu_native.rs:110-117
 *                                            src/rust/src/jit.rs:512-604
 *   - reverse-mode gradient tape             src/rust/src/cpu_native.rs:148-298
 *   - the whole-chain sampler `run`          src/rust/src/cpu_native.rs:510-825
 *     (modes 0 RWM, 1 DE path A, 2 Cholesky RWM, 3 MALA)
 *   - synchronous DE path B `run_de_sync`    src/rust/src/cpu_native.rs:400-506
 *   - loglik / pointwise / gradient batches  src/rust/src/cpu_native.rs:306-391
 *
 * Uniform streams.  The reference draws u = (triple32(seedmix + ctr) + 0.5)
 * / 2^32 with a pre-incremented per-chain counter.  Every mode consumes a
 * fixed number of uniforms per iteration (2*np + 1 for modes 0, 2, 3 and
 * 2*np + 4 for mode 1), so the counter at iteration t, slot s is
 * t * slots_per_iter + s + 1.  The oracle exposes that stream (STREAM_T32,
 * bit-for-bit the reference's) and the HIP backend's counter-based
 * Philox4x32-10 stream (STREAM_PHILOX) addressed by (t, chain, slot): the
 * same slot layout, a different generator.
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ---------------------------------------------------------------- RNG --- */

/* cpu_native.rs:13-23 */
uint32_t orc_triple32(uint32_t h) {
  h ^= h >> 17;
  h *= 2129725213u;
  h ^= h >> 11;
  h *= 2890025505u;
  h ^= h >> 15;
  h *= 831628273u;
  h ^= h >> 14;
  return h;
}

/* cpu_native.rs:25-28 */
double orc_uniform_t32(uint32_t seedmix, uint32_t ctr) {
  return ((double)orc_triple32(seedmix + ctr) + 0.5) / 4294967296.0;
}

/* cpu_native.rs:561-562 (also :436-438) */
uint32_t orc_seedmix(uint32_t seed, uint32_t chain) {
  return orc_triple32(seed) + chain * 2654435761u;
}

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11, "Parallel random
 * numbers: as easy as 1, 2, 3").  Multipliers 0xD2511F53 / 0xCD9E8D57, Weyl
 * key increments 0x9E3779B9 / 0xBB67AE85, ten rounds.  Not in the reference:
 * this is the HIP backend's generator, restated here so the oracle can
 * replay any uniform the kernel drew. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2],
                       uint32_t out[4]) {
  uint32_t x0 = ctr[0], x1 = ctr[1], x2 = ctr[2], x3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    uint64_t p0 = (uint64_t)0xD2511F53u * x0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * x2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t y0 = hi1 ^ x1 ^ k0;
    uint32_t y1 = lo1;
    uint32_t y2 = hi0 ^ x3 ^ k1;
    uint32_t y3 = lo0;
    x0 = y0; x1 = y1; x2 = y2; x3 = y3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

/* Uniform number `slot` of iteration `t` of chain `chain`, Philox stream:
 * counter (t, chain, slot / 4, tag), key (seed, stream), word slot % 4,
 * mapped to (h + 0.5) / 2^32 exactly like the reference's uniform. */
double orc_uniform_philox(uint32_t seed, uint32_t stream, uint32_t tag,
                          uint32_t t, uint32_t chain, uint32_t slot) {
  uint32_t ctr[4] = {t, chain, slot >> 2, tag};
  uint32_t key[2] = {seed, stream};
  uint32_t out[4];
  orc_philox4x32_10(ctr, key, out);
  return ((double)out[slot & 3u] + 0.5) / 4294967296.0;
}

/* ------------------------------------------------------------ special --- */

/* cpu_native.rs:34-60: Lanczos g = 7, nine coefficients, reflection for
 * x < 0.5. */
double orc_lgamma(double x) {
  static const double C[9] = {
      0.99999999999980993,     676.5203681218851,
      -1259.1392167224028,     771.32342877765313,
      -176.61502916214059,     12.507343278686905,
      -0.13857109526572012,    9.9843695780195716e-6,
      1.5056327351493116e-7};
  const double pi = 3.14159265358979323846;
  if (x < 0.5) {
    return log(fabs(pi / sin(pi * x))) - orc_lgamma(1.0 - x);
  }
  x = x - 1.0;
  double a = C[0];
  double t = x + 7.0 + 0.5;
  for (int i = 1; i < 9; ++i) a += C[i] / (x + (double)i);
  return 0.5 * log(2.0 * pi) + (x + 0.5) * log(t) - t + log(a);
}

/* cpu_native.rs:66-77 */
double orc_digamma(double x) {
  double result = 0.0;
  while (x < 6.0) {
    result -= 1.0 / x;
    x += 1.0;
  }
  double f = 1.0 / (x * x);
  return result + log(x) - 1.0 / (2.0 * x) -
         f * (1.0 / 12.0 - f * (1.0 / 120.0 - f * (1.0 / 252.0)));
}

/* -------------------------------------------------------- interpreter --- */

/* cpu_native.rs:80-107.  `code` holds n_words u32 words (op, arg pairs).
 * An empty program returns 0; an opcode outside {0..10, 12} executes as
 * POW, as in the reference's match default. */
double orc_vm_eval(const uint32_t* code, uint32_t n_words,
                   const double* consts, const double* params,
                   const double* row) {
  if (n_words < 2) return 0.0;
  double stack[32];
  int sp = 0;
  uint32_t n_instr = n_words / 2;
  for (uint32_t pc = 0; pc < n_instr; ++pc) {
    uint32_t op = code[2 * pc];
    uint32_t arg = code[2 * pc + 1];
    switch (op) {
      case 0: stack[sp++] = consts[arg]; break;
      case 1: stack[sp++] = params[arg]; break;
      case 2: stack[sp++] = row[arg]; break;
      case 3: stack[sp - 2] += stack[sp - 1]; --sp; break;
      case 4: stack[sp - 2] -= stack[sp - 1]; --sp; break;
      case 5: stack[sp - 2] *= stack[sp - 1]; --sp; break;
      case 6: stack[sp - 2] /= stack[sp - 1]; --sp; break;
      case 7: stack[sp - 1] = -stack[sp - 1]; break;
      case 8: stack[sp - 1] = exp(stack[sp - 1]); break;
      case 9: stack[sp - 1] = log(stack[sp - 1]); break;
      case 10: stack[sp - 1] = sqrt(stack[sp - 1]); break;
      case 12: stack[sp - 1] = orc_lgamma(stack[sp - 1]); break;
      default:
        stack[sp - 2] = pow(stack[sp - 2], stack[sp - 1]);
        --sp;
        break;
    }
  }
  return stack[0];
}

/* cpu_native.rs:110-117 (and the JIT's loop jit.rs:512-604): acc += term in
 * observation order. */
double orc_loglik(const uint32_t* code, uint32_t n_words, const double* consts,
                  const double* params, const double* data, uint32_t n_cols,
                  uint32_t n_obs) {
  double acc = 0.0;
  for (uint32_t i = 0; i < n_obs; ++i)
    acc += orc_vm_eval(code, n_words, consts, params,
                       data + (size_t)i * n_cols);
  return acc;
}

/* cpu_native.rs:148-279: reverse-mode tape of one program at one
 * observation; adds d(value)/d(params) into grad, returns the value. */
double orc_vm_eval_grad(const uint32_t* code, uint32_t n_words,
                        const double* consts, const double* params,
                        const double* row, double* vals, double* adj,
                        uint32_t* lhs, uint32_t* rhs, double* grad) {
  if (n_words < 2) return 0.0;
  uint32_t n_ops = n_words / 2;
  uint32_t istack[32];
  int sp = 0;
  const uint32_t NONE = 0xFFFFFFFFu;
  for (uint32_t pc = 0; pc < n_ops; ++pc) {
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    if (op <= 2) {
      vals[pc] = op == 0 ? consts[arg] : op == 1 ? params[arg] : row[arg];
      lhs[pc] = NONE;
      istack[sp++] = pc;
    } else if ((op >= 7 && op <= 10) || op == 12) {
      uint32_t l = istack[sp - 1];
      double x = vals[l];
      vals[pc] = op == 7 ? -x : op == 8 ? exp(x) : op == 9 ? log(x)
               : op == 10 ? sqrt(x) : orc_lgamma(x);
      lhs[pc] = l;
      rhs[pc] = NONE;
      istack[sp - 1] = pc;
    } else {
      uint32_t r = istack[sp - 1], l = istack[sp - 2];
      double a = vals[l], b = vals[r];
      vals[pc] = op == 3 ? a + b : op == 4 ? a - b : op == 5 ? a * b
               : op == 6 ? a / b : pow(a, b);
      lhs[pc] = l;
      rhs[pc] = r;
      istack[sp - 2] = pc;
      --sp;
    }
  }
  uint32_t root = n_ops - 1;
  for (uint32_t i = 0; i < n_ops; ++i) adj[i] = 0.0;
  adj[root] = 1.0;
  for (int64_t pc = (int64_t)n_ops - 1; pc >= 0; --pc) {
    double g = adj[pc];
    if (g == 0.0) continue;
    uint32_t op = code[2 * pc], arg = code[2 * pc + 1];
    switch (op) {
      case 0: case 2: break;
      case 1: grad[arg] += g; break;
      case 3: adj[lhs[pc]] += g; adj[rhs[pc]] += g; break;
      case 4: adj[lhs[pc]] += g; adj[rhs[pc]] -= g; break;
      case 5:
        adj[lhs[pc]] += g * vals[rhs[pc]];
        adj[rhs[pc]] += g * vals[lhs[pc]];
        break;
      case 6:
        adj[lhs[pc]] += g / vals[rhs[pc]];
        adj[rhs[pc]] -= g * vals[pc] / vals[rhs[pc]];
        break;
      case 7: adj[lhs[pc]] -= g; break;
      case 8: adj[lhs[pc]] += g * vals[pc]; break;
      case 9: adj[lhs[pc]] += g / vals[lhs[pc]]; break;
      case 10: adj[lhs[pc]] += g * 0.5 / vals[pc]; break;
      case 12: adj[lhs[pc]] += g * orc_digamma(vals[lhs[pc]]); break;
      default: {
        double a = vals[lhs[pc]], b = vals[rhs[pc]];
        adj[lhs[pc]] += g * b * pow(a, b - 1.0);
        if (a > 0.0) adj[rhs[pc]] += g * vals[pc] * log(a);
        break;
      }
    }
  }
  return vals[root];
}

typedef struct {
  double* vals;
  double* adj;
  uint32_t* lhs;
  uint32_t* rhs;
} grad_work;

static void gw_init(grad_work* w, uint32_t n_ops) {
  if (n_ops < 1) n_ops = 1;
  w->vals = (double*)calloc(n_ops, sizeof(double));
  w->adj = (double*)calloc(n_ops, sizeof(double));
  w->lhs = (uint32_t*)calloc(n_ops, sizeof(uint32_t));
  w->rhs = (uint32_t*)calloc(n_ops, sizeof(uint32_t));
}

static void gw_free(grad_work* w) {
  free(w->vals); free(w->adj); free(w->lhs); free(w->rhs);
}

/* ------------------------------------------------------------ batches --- */

/* cpu_native.rs:339-360 */
void orc_loglik_batch(const uint32_t* code, uint32_t n_words,
                      const double* consts, uint32_t n_params,
                      const double* data, uint32_t n_cols, uint32_t n_obs,
                      const double* points, uint32_t n_points, double* out) {
  for (uint32_t p = 0; p < n_points; ++p)
    out[p] = orc_loglik(code, n_words, consts, points + (size_t)p * n_params,
                        data, n_cols, n_obs);
}

/* cpu_native.rs:371-391: out[p * n_obs + i] */
void orc_loglik_pointwise(const uint32_t* code, uint32_t n_words,
                          const double* consts, uint32_t n_params,
                          const double* data, uint32_t n_cols, uint32_t n_obs,
                          const double* points, uint32_t n_points,
                          double* out) {
  for (uint32_t p = 0; p < n_points; ++p)
    for (uint32_t i = 0; i < n_obs; ++i)
      out[(size_t)p * n_obs + i] =
          orc_vm_eval(code, n_words, consts, points + (size_t)p * n_params,
                      data + (size_t)i * n_cols);
}

/* cpu_native.rs:306-329: gradient of the data term, point-major. */
void orc_grad_batch(const uint32_t* code, uint32_t n_words,
                    const double* consts, uint32_t n_params,
                    const double* data, uint32_t n_cols, uint32_t n_obs,
                    const double* points, uint32_t n_points, double* out) {
  grad_work w;
  gw_init(&w, n_words / 2);
  for (uint32_t p = 0; p < n_points; ++p) {
    double* g = out + (size_t)p * n_params;
    for (uint32_t j = 0; j < n_params; ++j) g[j] = 0.0;
    for (uint32_t i = 0; i < n_obs; ++i)
      orc_vm_eval_grad(code, n_words, consts, points + (size_t)p * n_params,
                       data + (size_t)i * n_cols, w.vals, w.adj, w.lhs, w.rhs,
                       g);
  }
  gw_free(&w);
}

/* ------------------------------------------------------------ sampler --- */

static double uniform_at(const orc_run_args* a, uint32_t chain, uint32_t t,
                         uint32_t slot, uint32_t spi) {
  if (a->stream == ORC_STREAM_T32) {
    uint32_t seedmix = orc_seedmix(a->seed, chain);
    return orc_uniform_t32(seedmix, t * spi + slot + 1u);
  }
  return orc_uniform_philox(a->seed, a->stream_id, 0u, a->iter_offset + t,
                            chain, slot);
}

static double log_post(const orc_run_args* a, const double* params) {
  double ll;
  if (a->loglik_fn)
    ll = a->loglik_fn(params, a->data, a->n_obs, a->n_cols);
  else
    ll = orc_loglik(a->code, a->code_len, a->consts, params, a->data,
                    a->n_cols, a->n_obs);
  return ll + orc_vm_eval(a->prior_code, a->prior_len, a->prior_consts,
                          params, NULL);
}

/* Value and gradient of the log-posterior: data term plus prior term
 * through the tape (cpu_native.rs:285-298). */
static double lp_grad(const orc_run_args* a, const double* params, double* g,
                      grad_work* w) {
  uint32_t np = a->n_params;
  for (uint32_t j = 0; j < np; ++j) g[j] = 0.0;
  double acc = 0.0;
  for (uint32_t i = 0; i < a->n_obs; ++i)
    acc += orc_vm_eval_grad(a->code, a->code_len, a->consts, params,
                            a->data + (size_t)i * a->n_cols, w->vals, w->adj,
                            w->lhs, w->rhs, g);
  return acc + orc_vm_eval_grad(a->prior_code, a->prior_len, a->prior_consts,
                                params, NULL, w->vals, w->adj, w->lhs, w->rhs,
                                g);
}

/* One chain of cpu_native.rs:555-807. */
static void run_chain(const orc_run_args* a, orc_run_out* o, uint32_t c) {
  const uint32_t np = a->n_params, C = a->n_chains, T = a->n_iter;
  const double two_pi = 6.283185307179586;
  const uint32_t mode = a->proposal_mode;
  const uint32_t spi = mode == 1 ? 2 * np + 4 : 2 * np + 1;
  const size_t pbase = (size_t)c * np;
  const size_t lbase = (size_t)c * np * np;
  const double temperature = a->temperatures[c];
  const int have_l = a->proposal_l != NULL;
  double state[ORC_MAX_PARAMS], prop[ORC_MAX_PARAMS], z[ORC_MAX_PARAMS];
  double gx[ORC_MAX_PARAMS], gy[ORC_MAX_PARAMS];
  double mean_x[ORC_MAX_PARAMS], mean_y[ORC_MAX_PARAMS];
  double vbuf[ORC_MAX_PARAMS];
  memcpy(state, a->init + pbase, np * sizeof(double));
  memcpy(prop, state, np * sizeof(double));
  grad_work w;
  int mala = mode == 3;
  uint32_t n_ops = a->code_len / 2;
  if (a->prior_len / 2 > n_ops) n_ops = a->prior_len / 2;
  gw_init(&w, n_ops);
  double cur = mala ? lp_grad(a, state, gx, &w) : log_post(a, state);
  uint32_t accepts = 0;
  const double drift_max = 2.0 * sqrt((double)np);

  for (uint32_t t = 0; t < T; ++t) {
    uint32_t s = 0; /* slot within the iteration */
    double plp;
    if (mala) {
      /* cpu_native.rs:606-724: preconditioned truncated Langevin step. */
      for (int pass = 0; pass < 1; ++pass) {
        /* half_m_times(gx) -> mean_x (cpu_native.rs:618-651) */
        if (!have_l) {
          for (uint32_t j = 0; j < np; ++j)
            vbuf[j] = 0.5 * a->proposal_sd[pbase + j] * gx[j];
        } else {
          for (uint32_t j = 0; j < np; ++j) {
            double acc = 0.0;
            for (uint32_t k = j; k < np; ++k)
              acc += a->proposal_l[lbase + k * np + j] * gx[k];
            vbuf[j] = 0.5 * acc;
          }
        }
        double nrm = 0.0;
        for (uint32_t j = 0; j < np; ++j) nrm += vbuf[j] * vbuf[j];
        nrm = sqrt(nrm);
        double f = nrm > drift_max ? drift_max / nrm : 1.0;
        if (!have_l) {
          for (uint32_t j = 0; j < np; ++j)
            mean_x[j] = f * a->proposal_sd[pbase + j] * vbuf[j];
        } else {
          for (uint32_t k = 0; k < np; ++k) {
            double acc = 0.0;
            for (uint32_t j = 0; j <= k; ++j)
              acc += a->proposal_l[lbase + k * np + j] * vbuf[j];
            mean_x[k] = f * acc;
          }
        }
      }
      for (uint32_t j = 0; j < np; ++j) {
        mean_x[j] += state[j];
        double u1 = uniform_at(a, c, t, s++, spi);
        double u2 = uniform_at(a, c, t, s++, spi);
        z[j] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
      }
      if (!have_l) {
        for (uint32_t j = 0; j < np; ++j)
          prop[j] = mean_x[j] + a->proposal_sd[pbase + j] * z[j];
      } else {
        for (uint32_t k = 0; k < np; ++k) {
          double acc = mean_x[k];
          for (uint32_t j = 0; j <= k; ++j)
            acc += a->proposal_l[lbase + k * np + j] * z[j];
          prop[k] = acc;
        }
      }
      plp = lp_grad(a, prop, gy, &w);
      /* half_m_times(gy) -> mean_y */
      if (!have_l) {
        for (uint32_t j = 0; j < np; ++j)
          vbuf[j] = 0.5 * a->proposal_sd[pbase + j] * gy[j];
      } else {
        for (uint32_t j = 0; j < np; ++j) {
          double acc = 0.0;
          for (uint32_t k = j; k < np; ++k)
            acc += a->proposal_l[lbase + k * np + j] * gy[k];
          vbuf[j] = 0.5 * acc;
        }
      }
      {
        double nrm = 0.0;
        for (uint32_t j = 0; j < np; ++j) nrm += vbuf[j] * vbuf[j];
        nrm = sqrt(nrm);
        double f = nrm > drift_max ? drift_max / nrm : 1.0;
        if (!have_l) {
          for (uint32_t j = 0; j < np; ++j)
            mean_y[j] = f * a->proposal_sd[pbase + j] * vbuf[j];
        } else {
          for (uint32_t k = 0; k < np; ++k) {
            double acc = 0.0;
            for (uint32_t j = 0; j <= k; ++j)
              acc += a->proposal_l[lbase + k * np + j] * vbuf[j];
            mean_y[k] = f * acc;
          }
        }
      }
      for (uint32_t j = 0; j < np; ++j) mean_y[j] += prop[j];
      /* log_q (cpu_native.rs:655-673) */
      double q1 = 0.0, q2 = 0.0;
      const double* vs[2] = {state, prop};
      const double* ms[2] = {mean_y, mean_x};
      for (int qq = 0; qq < 2; ++qq) {
        double q = 0.0;
        const double* v = vs[qq];
        const double* m = ms[qq];
        if (!have_l) {
          for (uint32_t j = 0; j < np; ++j) {
            double r = (v[j] - m[j]) / a->proposal_sd[pbase + j];
            q -= 0.5 * r * r;
          }
        } else {
          for (uint32_t k = 0; k < np; ++k) {
            double acc = v[k] - m[k];
            for (uint32_t j = 0; j < k; ++j)
              acc -= a->proposal_l[lbase + k * np + j] * vbuf[j];
            vbuf[k] = acc / a->proposal_l[lbase + k * np + k];
            q -= 0.5 * vbuf[k] * vbuf[k];
          }
        }
        if (qq == 0) q1 = q; else q2 = q;
      }
      double logq_diff = q1 - q2;
      double u = uniform_at(a, c, t, s++, spi);
      int acc_ok = isfinite(plp) && log(u) < (plp - cur) / temperature + logq_diff;
      if (o->trace_plp) o->trace_plp[(size_t)t * C + c] = plp;
      if (o->trace_prop)
        memcpy(o->trace_prop + ((size_t)t * C + c) * np, prop,
               np * sizeof(double));
      if (acc_ok) {
        memcpy(state, prop, np * sizeof(double));
        memcpy(gx, gy, np * sizeof(double));
        cur = plp;
        ++accepts;
      }
      if (o->trace_accept) o->trace_accept[(size_t)t * C + c] = (uint8_t)acc_ok;
      for (uint32_t j = 0; j < np; ++j)
        o->draws[t + (size_t)T * (c + (size_t)C * j)] = (float)state[j];
      continue;
    }
    if (mode == 0) {
      /* cpu_native.rs:726-735 */
      for (uint32_t j = 0; j < np; ++j) {
        double u1 = uniform_at(a, c, t, s++, spi);
        double u2 = uniform_at(a, c, t, s++, spi);
        double zz = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
        prop[j] = state[j] + a->proposal_sd[pbase + j] * zz;
      }
    } else if (mode == 2) {
      /* cpu_native.rs:736-753 */
      for (uint32_t j = 0; j < np; ++j) {
        double u1 = uniform_at(a, c, t, s++, spi);
        double u2 = uniform_at(a, c, t, s++, spi);
        z[j] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
      }
      for (uint32_t k = 0; k < np; ++k) {
        double acc = state[k];
        for (uint32_t j = 0; j <= k; ++j)
          acc += a->proposal_l[lbase + k * np + j] * z[j];
        prop[k] = acc;
      }
    } else {
      /* cpu_native.rs:754-784: DE path A against the frozen `init`. */
      double ua = uniform_at(a, c, t, s++, spi);
      uint32_t ia = (uint32_t)(ua * (double)C);
      if (ia >= C) ia = C - 1;
      double ub = uniform_at(a, c, t, s++, spi);
      uint32_t ib = (uint32_t)(ub * (double)C);
      if (ib >= C) ib = C - 1;
      if (ib == ia) ib = (ib + 1) % C;
      double ug = uniform_at(a, c, t, s++, spi);
      double g = ug < 0.1 ? 1.0 : a->gamma;
      for (uint32_t j = 0; j < np; ++j) {
        double u1 = uniform_at(a, c, t, s++, spi);
        double u2 = uniform_at(a, c, t, s++, spi);
        double zz = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
        double diff = a->init[(size_t)ia * np + j] - a->init[(size_t)ib * np + j];
        prop[j] = state[j] + g * diff + a->de_noise * a->proposal_sd[pbase + j] * zz;
      }
    }
    plp = log_post(a, prop);
    /* cpu_native.rs:785-796 */
    double u = uniform_at(a, c, t, s++, spi);
    int acc_ok = log(u) < (plp - cur) / temperature;
    if (o->trace_plp) o->trace_plp[(size_t)t * C + c] = plp;
    if (o->trace_prop)
      memcpy(o->trace_prop + ((size_t)t * C + c) * np, prop,
             np * sizeof(double));
    if (acc_ok) {
      memcpy(state, prop, np * sizeof(double));
      cur = plp;
      ++accepts;
    }
    if (o->trace_accept) o->trace_accept[(size_t)t * C + c] = (uint8_t)acc_ok;
    /* cpu_native.rs:797-799 and the column-major assembly :810-822 */
    for (uint32_t j = 0; j < np; ++j)
      o->draws[t + (size_t)T * (c + (size_t)C * j)] = (float)state[j];
  }
  o->accept_rate[c] = T == 0 ? 0.0f : (float)accepts / (float)T;
  o->last_log_post[c] = cur;
  if (o->final_state) memcpy(o->final_state + pbase, state, np * sizeof(double));
  gw_free(&w);
}

typedef struct {
  const orc_run_args* a;
  orc_run_out* o;
  uint32_t next;
  pthread_mutex_t mu;
} pool_ctx;

static void* pool_worker(void* p) {
  pool_ctx* ctx = (pool_ctx*)p;
  for (;;) {
    pthread_mutex_lock(&ctx->mu);
    uint32_t c = ctx->next++;
    pthread_mutex_unlock(&ctx->mu);
    if (c >= ctx->a->n_chains) break;
    run_chain(ctx->a, ctx->o, c);
  }
  return NULL;
}

/* cpu_native.rs:510-825, parallel over chains (rayon there, a pthread pool
 * here).  Returns 0 on success, nonzero on invalid arguments. */
int orc_run(const orc_run_args* a, orc_run_out* o) {
  if (a->n_params == 0 || a->n_params > ORC_MAX_PARAMS) return 1;
  if (a->proposal_mode > 3) return 2;
  if (a->proposal_mode == 2 && a->proposal_l == NULL) return 3;
  uint32_t nt = a->n_threads ? a->n_threads : 1;
  if (nt > a->n_chains) nt = a->n_chains ? a->n_chains : 1;
  pool_ctx ctx;
  ctx.a = a;
  ctx.o = o;
  ctx.next = 0;
  pthread_mutex_init(&ctx.mu, NULL);
  if (nt <= 1) {
    pool_worker(&ctx);
  } else {
    pthread_t* th = (pthread_t*)malloc(nt * sizeof(pthread_t));
    for (uint32_t i = 0; i < nt; ++i)
      pthread_create(&th[i], NULL, pool_worker, &ctx);
    for (uint32_t i = 0; i < nt; ++i) pthread_join(th[i], NULL);
    free(th);
  }
  pthread_mutex_destroy(&ctx.mu);
  return 0;
}

/* cpu_native.rs:400-506: synchronous DE (path B), one generation at a time,
 * the pair excludes the proposing chain.  Uses the reference's running
 * counter (2*np + 4 uniforms per chain per generation). */
int orc_run_de_sync(const orc_run_args* a, orc_run_out* o) {
  const uint32_t np = a->n_params, C = a->n_chains, T = a->n_iter;
  if (np == 0 || np > ORC_MAX_PARAMS || C < 3) return 1;
  const double two_pi = 6.283185307179586;
  const uint32_t spi = 2 * np + 4;
  double* cur_pop = (double*)malloc((size_t)C * np * sizeof(double));
  double* next_pop = (double*)malloc((size_t)C * np * sizeof(double));
  double* cur_lp = (double*)malloc((size_t)C * sizeof(double));
  uint32_t* accepts = (uint32_t*)calloc(C, sizeof(uint32_t));
  memcpy(cur_pop, a->init, (size_t)C * np * sizeof(double));
  for (uint32_t c = 0; c < C; ++c) cur_lp[c] = log_post(a, cur_pop + (size_t)c * np);
  double prop[ORC_MAX_PARAMS];
  for (uint32_t t = 0; t < T; ++t) {
    for (uint32_t c = 0; c < C; ++c) {
      uint32_t s = 0;
      const double* st = cur_pop + (size_t)c * np;
      uint32_t ia = (uint32_t)(uniform_at(a, c, t, s++, spi) * (double)C);
      if (ia >= C) ia = C - 1;
      if (ia == c) ia = (ia + 1) % C;
      uint32_t ib = (uint32_t)(uniform_at(a, c, t, s++, spi) * (double)C);
      if (ib >= C) ib = C - 1;
      while (ib == c || ib == ia) ib = (ib + 1) % C;
      double g = uniform_at(a, c, t, s++, spi) < 0.1 ? 1.0 : a->gamma;
      for (uint32_t j = 0; j < np; ++j) {
        double u1 = uniform_at(a, c, t, s++, spi);
        double u2 = uniform_at(a, c, t, s++, spi);
        double zz = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
        double diff = cur_pop[(size_t)ia * np + j] - cur_pop[(size_t)ib * np + j];
        prop[j] = st[j] + g * diff + a->de_noise * a->proposal_sd[(size_t)c * np + j] * zz;
      }
      double plp = log_post(a, prop);
      double u = uniform_at(a, c, t, s++, spi);
      double* nx = next_pop + (size_t)c * np;
      const int acc = log(u) < plp - cur_lp[c];
      if (o->trace_plp) o->trace_plp[(size_t)t * C + c] = plp;
      if (o->trace_prop)
        for (uint32_t j = 0; j < np; ++j)
          o->trace_prop[((size_t)t * C + c) * np + j] = prop[j];
      if (o->trace_accept) o->trace_accept[(size_t)t * C + c] = (uint8_t)acc;
      if (acc) {
        memcpy(nx, prop, np * sizeof(double));
        cur_lp[c] = plp;
        ++accepts[c];
      } else {
        memcpy(nx, st, np * sizeof(double));
      }
      for (uint32_t j = 0; j < np; ++j)
        o->draws[t + (size_t)T * (c + (size_t)C * j)] = (float)nx[j];
    }
    double* tmp = cur_pop; cur_pop = next_pop; next_pop = tmp;
  }
  for (uint32_t c = 0; c < C; ++c) {
    o->accept_rate[c] = T == 0 ? 0.0f : (float)accepts[c] / (float)T;
    o->last_log_post[c] = cur_lp[c];
  }
  if (o->final_state) memcpy(o->final_state, cur_pop, (size_t)C * np * sizeof(double));
  free(cur_pop); free(next_pop); free(cur_lp); free(accepts);
  return 0;
}
