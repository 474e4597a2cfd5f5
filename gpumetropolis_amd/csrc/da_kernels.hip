// da_kernels.hip -- the stepwise sampler's per-iteration kernels.
//
// One iteration is a short sequence of launches on one stream:
//   propose   one wave per chain, a lane per parameter: Philox normals and
//             the proposal (cpu_native.rs:726-784, modes 0 and 2), then its
//             prior (lead lane) and with a surrogate the network's input
//             row (a lane per input);
//   [network forward over all proposals]
//   stage1    one workgroup: the surrogate's Metropolis test
//             log u1 < (s(prop) - s(theta)) / T, and an ordered compaction of
//             the surviving proposals into the exact stage's slot list
//             (without a surrogate the propose kernel writes the list, every
//             chain in order, and this launch is skipped);
//   points    the exact data term of the listed proposals (chain kernel
//             file), split over observation slices;
//   stage2    one thread per chain: the exact test, corrected by the
//             surrogate ratio so the target is unchanged (Christen and Fox,
//             "Markov chain Monte Carlo using an approximation", JCGS 2005):
//             log u < ((lp' - lp) - (s' - s)) / T; record the draw in an
//             iteration-major staging buffer (coalesced), turned into the
//             reference's layout once after the run.
// Without a surrogate every proposal passes stage 1 and stage 2 is the
// reference's Metropolis test (cpu_native.rs:785-799) with the same
// uniforms as the fused chain kernel.  Mode 1 (synchronous Differential
// Evolution) changes only the proposal.  Mode 3 (Langevin) runs its own
// propose and accept kernels around the FP64 value-and-gradient kernel of
// eval64.hip.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "bf16.h"
#include "da.h"
#include "mlp.h"
#include "philox.h"
#include "step_common.h"

namespace gpum {

__device__ __forceinline__ void write_features(const double* __restrict__ p,
                                               uint32_t n_in,
                                               const double* __restrict__ mean,
                                               const double* __restrict__ w,
                                               uint16_t* __restrict__ row) {
  for (uint32_t k = 0; k < (uint32_t)kMlpIn; ++k) {
    double x = 0.0;
    if (k < n_in)
      for (uint32_t j = 0; j < n_in; ++j) x += w[k * n_in + j] * (p[j] - mean[j]);
    row[k] = f2bf((float)x);
  }
}

__global__ void da_features_kernel(const double* __restrict__ points, uint32_t n,
                                   uint32_t n_in, const double* __restrict__ mean,
                                   const double* __restrict__ w,
                                   uint16_t* __restrict__ Xb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  write_features(points + (size_t)i * n_in, n_in, mean, w, Xb + (size_t)i * kMlpIn);
}

__global__ void da_init_kernel(const DaArgs* __restrict__ A) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A->n_chains) return;
  const uint32_t np = A->n_params;
  const double* th = A->theta + (size_t)c * np;
  A->prop_prior[c] = eval_params64(A->prior, A->n_prior, A->prior_consts, th, nullptr);
  A->accepts[c] = 0.0;
  A->passes[c] = 0.0;
  if (A->use_surrogate)
    write_features(th, A->n_in, A->norm_mean, A->norm_w, A->Xb + (size_t)c * kMlpIn);
}

// cur = data term of the initial state (slot c) + its prior; the reference
// evaluates the start before its loop (cpu_native.rs:565).
__global__ void da_init_finish_kernel(const DaArgs* __restrict__ A) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C = A->n_chains;
  if (c >= C) return;
  double ll = 0.0;
  const uint32_t ns = *A->n_slices;
  for (uint32_t s = 0; s < ns; ++s) ll += A->partial[(size_t)s * C + c];
  const double pr = A->prop_prior[c];
  A->cur[c] = ll + pr;
  A->cur_s[c] = A->use_surrogate
                    ? pr + (A->t_mean + A->t_sd * (double)A->s_raw[c])
                    : 0.0;
}

constexpr int kProposeWaves = 4;  // chains per 256-thread workgroup

__global__ void __launch_bounds__(256) da_propose_kernel(const DaArgs* __restrict__ A,
                                                         uint32_t t) {
  __shared__ double prop_sh[kProposeWaves][kMaxParams];
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const uint32_t c = blockIdx.x * kProposeWaves + (uint32_t)w;
  const bool live = c < A->n_chains;  // uniform across the wave
  const uint32_t np = A->n_params;
  const bool mine = live && (uint32_t)lane < np;  // lane j: parameter j
  if (live) {
    const double two_pi = 6.283185307179586;
    StepUniforms us{A->seed, A->stream_id, A->iter_offset + t, c};
    // Slots of the normal deviates: after a, b and the scale choice in
    // mode 1 (cpu_native.rs:455-471).
    const uint32_t zo = A->mode == 1 ? 3u : 0u;
    double z = 0.0;
    if (mine) {
      const double u1 = us.get(zo + 2 * lane), u2 = us.get(zo + 2 * lane + 1);
      z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
    }
    const double* th = A->theta + (size_t)c * np;
    double p = 0.0;
    if (A->mode == 1) {
      // Rounded operation by operation as the reference (no fused
      // multiply-add): each generation's states are sums of the previous
      // generation's differences, so a last-place difference would grow
      // from generation to generation.
#pragma clang fp contract(off)
      // The pair (a, b) from the current generation, a != c, b not in
      // {c, a} (the host requires at least 3 chains); with probability 0.1
      // the scale is 1.  Every lane of the wave draws the same pair.
      const uint32_t C = A->n_chains;
      uint32_t ia = (uint32_t)(us.get(0) * (double)C);
      if (ia >= C) ia = C - 1;
      if (ia == c) ia = (ia + 1) % C;
      uint32_t ib = (uint32_t)(us.get(1) * (double)C);
      if (ib >= C) ib = C - 1;
      while (ib == c || ib == ia) ib = (ib + 1) % C;
      const double g = us.get(2) < 0.1 ? 1.0 : A->gamma;
      if (mine) {
        const double diff = A->theta[(size_t)ia * np + lane] -
                            A->theta[(size_t)ib * np + lane];
        p = th[lane] + g * diff +
            A->de_noise * A->proposal_sd[(size_t)c * np + lane] * z;
      }
    } else if (A->mode == 2) {
      // pr[k] = th[k] + sum_{j <= k} L[k][j] z[j], j ascending as the
      // reference sums it.
      const double* L = A->proposal_l + (size_t)c * np * np;
      double a = mine ? th[lane] : 0.0;
      for (uint32_t j = 0; j < np; ++j) {
        const double zj = __shfl(z, (int)j, 64);
        if (mine && j <= (uint32_t)lane) a += L[(uint32_t)lane * np + j] * zj;
      }
      p = a;
    } else if (mine) {
      p = th[lane] + A->proposal_sd[(size_t)c * np + lane] * z;
    }
    if (mine) {
      A->prop[(size_t)c * np + lane] = p;
      prop_sh[w][lane] = p;
    }
  }
  __syncthreads();
  if (!live) return;
  if (lane == 0) {
    A->prop_prior[c] = eval_params64(A->prior, A->n_prior, A->prior_consts,
                                     prop_sh[w], nullptr);
    if (!A->use_surrogate) {
      // Every proposal goes to the exact stage, in chain order: the slot
      // list da_stage1_kernel would write, written here instead.
      A->list[c] = c;
      A->pos[c] = (int32_t)c;
      A->passes[c] += 1.0;
      if (c == 0) *A->count = A->n_chains;
    }
  }
  if (A->use_surrogate && lane < kMlpIn) {
    const uint32_t n_in = A->n_in;
    double x = 0.0;
    if ((uint32_t)lane < n_in)
      for (uint32_t j = 0; j < n_in; ++j)
        x += A->norm_w[(uint32_t)lane * n_in + j] * (prop_sh[w][j] - A->norm_mean[j]);
    A->Xb[(size_t)c * kMlpIn + lane] = f2bf((float)x);
  }
}

// One workgroup of 1024 threads walks the chains in order, so the slot of
// every surviving proposal is its rank among the survivors.
__global__ void __launch_bounds__(1024) da_stage1_kernel(const DaArgs* __restrict__ A,
                                                         uint32_t t) {
  __shared__ uint32_t wave_tot[16];
  const uint32_t C = A->n_chains;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  uint32_t base = 0;
  for (uint32_t c0 = 0; c0 < C; c0 += 1024) {
    const uint32_t c = c0 + (uint32_t)tid;
    bool pass = false;
    if (c < C) {
      if (A->use_surrogate) {
        const double sp = A->prop_prior[c] + (A->t_mean + A->t_sd * (double)A->s_raw[c]);
        A->prop_s[c] = sp;
        const U4 w = philox4x32_10(U4{A->iter_offset + t, c, 0u, TAG_DA}, A->seed,
                                   A->stream_id);
        const double u1 = u32_to_unit(w.x);
        pass = isfinite(sp) && log(u1) < (sp - A->cur_s[c]) / A->temperatures[c];
        if (A->trace_s) A->trace_s[(size_t)t * C + c] = sp;
      } else {
        pass = true;
      }
    }
    const unsigned long long mask = __ballot(pass);
    const uint32_t rank =
        (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) wave_tot[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t off = base, total = 0;
    for (int w = 0; w < 16; ++w) {
      if (w < wave) off += wave_tot[w];
      total += wave_tot[w];
    }
    if (c < C) {
      if (pass) {
        A->list[off + rank] = c;
        A->pos[c] = (int32_t)(off + rank);
        A->passes[c] += 1.0;
      } else {
        A->pos[c] = -1;
      }
    }
    base += total;
    __syncthreads();
  }
  if (tid == 0) *A->count = base;
}

__global__ void da_stage2_kernel(const DaArgs* __restrict__ A, uint32_t t) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t C = A->n_chains;
  if (c >= C) return;
  const uint32_t np = A->n_params;
  const int32_t p = A->pos[c];
  bool acc_ok = false;
  double lpp = __builtin_nan("");
  double ll = __builtin_nan("");
  if (p >= 0) {
    ll = 0.0;
    const uint32_t ns = *A->n_slices;
    const double* part = A->partial + p;
    // Slice order; eight slices' partials are loaded before they are added.
    uint32_t s = 0;
    for (; s + 8 <= ns; s += 8) {
      double v[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) v[k] = part[(size_t)(s + k) * C];
#pragma unroll
      for (uint32_t k = 0; k < 8; ++k) ll += v[k];
    }
    for (; s < ns; ++s) ll += part[(size_t)s * C];
    lpp = ll + A->prop_prior[c];
    double la = lpp - A->cur[c];
    if (A->use_surrogate) la -= A->prop_s[c] - A->cur_s[c];
    StepUniforms us{A->seed, A->stream_id, A->iter_offset + t, c};
    // The last slot of the iteration: 2 np + 1 draws, 2 np + 4 in mode 1.
    const double u = us.get(A->mode == 1 ? 2 * np + 3 : 2 * np);
    acc_ok = log(u) < la / A->temperatures[c];
  }
  double* th = A->theta + (size_t)c * np;
  if (acc_ok) {
    const double* pr = A->prop + (size_t)c * np;
    for (uint32_t j = 0; j < np; ++j) th[j] = pr[j];
    A->cur[c] = lpp;
    if (A->use_surrogate) A->cur_s[c] = A->prop_s[c];
    A->accepts[c] += 1.0;
  }
  if (A->trace_plp) A->trace_plp[(size_t)t * C + c] = lpp;
  if (A->trace_ll) A->trace_ll[(size_t)t * C + c] = ll;
  if (A->trace_prop)
    for (uint32_t j = 0; j < np; ++j)
      A->trace_prop[((size_t)t * C + c) * np + j] = A->prop[(size_t)c * np + j];
  if (A->trace_stage) A->trace_stage[(size_t)t * C + c] = p < 0 ? 0 : (acc_ok ? 2 : 1);
  for (uint32_t j = 0; j < np; ++j)
    A->draws_it[((size_t)t * np + j) * C + c] = (float)th[j];
}

// draws[t + n_iter * r] = draws_it[t * rows + r] with r = c + C j: a
// transpose of the (n_iter x rows) staging matrix through LDS tiles.
__global__ void __launch_bounds__(256) da_draws_kernel(const float* __restrict__ in,
                                                       float* __restrict__ out,
                                                       uint32_t n_iter, uint32_t rows) {
  __shared__ float tile[32][33];
  const uint32_t r0 = blockIdx.x * 32, t0 = blockIdx.y * 32;
  const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (uint32_t k = ty; k < 32; k += 8) {
    const uint32_t t = t0 + k, r = r0 + tx;
    if (t < n_iter && r < rows) tile[k][tx] = in[(size_t)t * rows + r];
  }
  __syncthreads();
  for (uint32_t k = ty; k < 32; k += 8) {
    const uint32_t r = r0 + k, t = t0 + tx;
    if (t < n_iter && r < rows) out[(size_t)r * n_iter + t] = tile[tx][k];
  }
}

// ---- Langevin steps (mode 3), cpu_native.rs:573-724 -------------------
// One thread per chain; n_params <= 16, so each chain's vectors are short.
// Every expression is rounded as the reference's (no fused multiply-add).

// The Langevin kernels run one wave per chain, lane k holding parameter k
// (k < d); every sum keeps the order in which the reference's loops add
// its terms.

// Lane j's x in every lane (j the same in every lane).
__device__ __forceinline__ double from_lane(double x, uint32_t j) {
  const uint64_t b = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = __builtin_amdgcn_readlane((uint32_t)b, (int)j);
  const uint32_t hi = __builtin_amdgcn_readlane((uint32_t)(b >> 32), (int)j);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// The preconditioned, truncated drift f * 0.5 * M g with M = Ls Ls' (Ls
// the chain's Cholesky factor, or the diagonal proposal_sd), the whitened
// drift norm capped at 2 sqrt(d) (cpu_native.rs:618-651): lane k's
// component from the gradient's component g in lane k.
__device__ double mala_drift(const DaArgs* __restrict__ A, uint32_t c,
                             double g, uint32_t lane) {
#pragma clang fp contract(off)
  const uint32_t np = A->n_params;
  const bool mine = lane < np;
  const double dmax = 2.0 * sqrt((double)np);
  const double* sd = A->proposal_sd + (size_t)c * np;
  const double* L = A->proposal_l ? A->proposal_l + (size_t)c * np * np : nullptr;
  // tmp_j = 0.5 * (Ls' g)_j, k ascending from j.
  double tmp = 0.0;
  if (!L) {
    if (mine) tmp = 0.5 * sd[lane] * g;
  } else {
    double acc = 0.0;
    for (uint32_t k = 0; k < np; ++k) {
      const double gk = from_lane(g, k);
      if (mine && k >= lane) acc += L[k * np + lane] * gk;
    }
    tmp = 0.5 * acc;
  }
  double nrm = 0.0;
  for (uint32_t j = 0; j < np; ++j) {
    const double tj = from_lane(tmp, j);
    nrm += tj * tj;
  }
  nrm = sqrt(nrm);
  const double f = nrm > dmax ? dmax / nrm : 1.0;
  if (!mine) return 0.0;
  if (!L) return f * sd[lane] * tmp;
  double acc = 0.0;
  for (uint32_t j = 0; j < np; ++j) {
    const double tj = from_lane(tmp, j);
    if (j <= lane) acc += L[lane * np + j] * tj;
  }
  return f * acc;
}

// -0.5 |Ls^{-1} (v - m)|^2 (cpu_native.rs:655-675) from lane k's
// components v and m.  The forward substitution runs by columns: once
// tmp_j is known every later row subtracts its term, so each row still
// subtracts in ascending j.
__device__ double mala_log_q(const DaArgs* __restrict__ A, uint32_t c,
                             double v, double m, uint32_t lane) {
#pragma clang fp contract(off)
  const uint32_t np = A->n_params;
  const bool mine = lane < np;
  const double* sd = A->proposal_sd + (size_t)c * np;
  const double* L = A->proposal_l ? A->proposal_l + (size_t)c * np * np : nullptr;
  double q = 0.0;
  if (!L) {
    const double r = mine ? (v - m) / sd[lane] : 0.0;
    for (uint32_t k = 0; k < np; ++k) {
      const double rk = from_lane(r, k);
      q -= 0.5 * rk * rk;
    }
    return q;
  }
  double acc = v - m;
  for (uint32_t k = 0; k < np; ++k) {
    const double tk_own = mine && lane == k ? acc / L[k * np + k] : 0.0;
    const double tk = from_lane(tk_own, k);
    q -= 0.5 * tk * tk;
    if (mine && lane > k) acc -= L[lane * np + k] * tk;
  }
  return q;
}

// mala_log_q(v1, m1) - mala_log_q(v2, m2): the two substitutions advance in
// one loop (each with its own operations in its own order), so their
// serial steps overlap.
__device__ double mala_log_q_diff(const DaArgs* __restrict__ A, uint32_t c,
                                  double v1, double m1, double v2, double m2,
                                  uint32_t lane) {
#pragma clang fp contract(off)
  const uint32_t np = A->n_params;
  const double* L = A->proposal_l ? A->proposal_l + (size_t)c * np * np : nullptr;
  if (!L)
    return mala_log_q(A, c, v1, m1, lane) - mala_log_q(A, c, v2, m2, lane);
  const bool mine = lane < np;
  double q1 = 0.0, q2 = 0.0;
  double acc1 = v1 - m1, acc2 = v2 - m2;
  for (uint32_t k = 0; k < np; ++k) {
    const double lkk = L[k * np + k];
    const bool own = mine && lane == k;
    const double t1 = from_lane(own ? acc1 / lkk : 0.0, k);
    const double t2 = from_lane(own ? acc2 / lkk : 0.0, k);
    q1 -= 0.5 * t1 * t1;
    q2 -= 0.5 * t2 * t2;
    if (mine && lane > k) {
      const double lk = L[lane * np + k];
      acc1 -= lk * t1;
      acc2 -= lk * t2;
    }
  }
  return q1 - q2;
}

// The log-posterior's gradient at a point: the data term's, then the
// prior's added (as the reference adds the prior's adjoints after the data
// loop, cpu_native.rs:285-298).
__device__ __forceinline__ void mala_gradient(const DaArgs* __restrict__ A,
                                              uint32_t c,
                                              double (&g)[kMaxParams]) {
#pragma clang fp contract(off)
  const uint32_t np = A->n_params;
  for (int k = 0; k < kMaxParams; ++k)
    g[k] = (uint32_t)k < np ? A->gy[(size_t)c * np + k] +
                                  A->g_prior[(size_t)c * np + k]
                            : 0.0;
}

__global__ void mala_init_kernel(const DaArgs* __restrict__ A) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= A->n_chains) return;
  const uint32_t np = A->n_params;
  double g[kMaxParams];
  mala_gradient(A, c, g);
  A->cur[c] = A->ll_prop[c] + A->v_prior[c];
  for (uint32_t k = 0; k < np; ++k) A->gx[(size_t)c * np + k] = g[k];
}

constexpr int kMalaWaves = 4;  // chains per 256-thread workgroup

__global__ void __launch_bounds__(256) mala_propose_kernel(
    const DaArgs* __restrict__ A, uint32_t t) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t c = blockIdx.x * kMalaWaves + (threadIdx.x >> 6);
  if (c >= A->n_chains) return;  // the whole wave
  const uint32_t np = A->n_params;
  const bool mine = lane < np;
  const double two_pi = 6.283185307179586;
  const double th = mine ? A->theta[(size_t)c * np + lane] : 0.0;
  const double gx = mine ? A->gx[(size_t)c * np + lane] : 0.0;
  const double mx = mala_drift(A, c, gx, lane) + th;
  StepUniforms us{A->seed, A->stream_id, A->iter_offset + t, c};
  double z = 0.0;
  if (mine) {
    const double u1 = us.get(2 * lane), u2 = us.get(2 * lane + 1);
    z = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
  }
  const double* sd = A->proposal_sd + (size_t)c * np;
  const double* L = A->proposal_l ? A->proposal_l + (size_t)c * np * np : nullptr;
  double p = mx;
  if (!L) {
    if (mine) p = mx + sd[lane] * z;
  } else {
    for (uint32_t j = 0; j < np; ++j) {
      const double zj = from_lane(z, j);
      if (mine && j <= lane) p += L[lane * np + j] * zj;
    }
  }
  if (mine) {
    A->prop[(size_t)c * np + lane] = p;
    A->mean_x[(size_t)c * np + lane] = mx;
  }
}

__global__ void __launch_bounds__(256) mala_accept_kernel(
    const DaArgs* __restrict__ A, uint32_t t) {
#pragma clang fp contract(off)
  const uint32_t lane = threadIdx.x & 63u;
  const uint32_t c = blockIdx.x * kMalaWaves + (threadIdx.x >> 6);
  const uint32_t C = A->n_chains;
  if (c >= C) return;  // the whole wave
  const uint32_t np = A->n_params;
  const bool mine = lane < np;
  const size_t ck = (size_t)c * np + lane;
  const double pr = mine ? A->prop[ck] : 0.0;
  const double th = mine ? A->theta[ck] : 0.0;
  const double mxo = mine ? A->mean_x[ck] : 0.0;
  const double g = mine ? A->gy[ck] + A->g_prior[ck] : 0.0;
  const double ll = A->ll_prop[c];
  const double plp = ll + A->v_prior[c];
  const double my = mala_drift(A, c, g, lane) + pr;
  const double lq = mala_log_q_diff(A, c, th, my, pr, mxo, lane);
  StepUniforms us{A->seed, A->stream_id, A->iter_offset + t, c};
  const double u = us.get(2 * np);
  const bool acc_ok =
      isfinite(plp) && log(u) < (plp - A->cur[c]) / A->temperatures[c] + lq;
  const double now = acc_ok ? pr : th;
  if (mine) {
    if (acc_ok) {
      A->theta[ck] = pr;
      A->gx[ck] = g;
    }
    if (A->trace_prop) A->trace_prop[((size_t)t * C + c) * np + lane] = pr;
    A->draws_it[((size_t)t * np + lane) * C + c] = (float)now;
  }
  if (lane == 0) {
    if (acc_ok) {
      A->cur[c] = plp;
      A->accepts[c] += 1.0;
    }
    A->passes[c] += 1.0;
    if (A->trace_plp) A->trace_plp[(size_t)t * C + c] = plp;
    if (A->trace_ll) A->trace_ll[(size_t)t * C + c] = ll;
    if (A->trace_stage) A->trace_stage[(size_t)t * C + c] = acc_ok ? 2 : 1;
  }
}

static dim3 chain_grid(uint32_t n) { return dim3((n + 255) / 256); }

hipError_t launch_mala_init(const DaArgs* A, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(mala_init_kernel, chain_grid(n), dim3(256), 0, s, A);
  return hipGetLastError();
}
hipError_t launch_mala_propose(const DaArgs* A, uint32_t n, uint32_t t,
                               hipStream_t s) {
  hipLaunchKernelGGL(mala_propose_kernel, dim3((n + kMalaWaves - 1) / kMalaWaves),
                     dim3(256), 0, s, A, t);
  return hipGetLastError();
}
hipError_t launch_mala_accept(const DaArgs* A, uint32_t n, uint32_t t,
                              hipStream_t s) {
  hipLaunchKernelGGL(mala_accept_kernel, dim3((n + kMalaWaves - 1) / kMalaWaves),
                     dim3(256), 0, s, A, t);
  return hipGetLastError();
}

hipError_t launch_da_init(const DaArgs* A, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(da_init_kernel, chain_grid(n), dim3(256), 0, s, A);
  return hipGetLastError();
}
hipError_t launch_da_init_finish(const DaArgs* A, uint32_t n, hipStream_t s) {
  hipLaunchKernelGGL(da_init_finish_kernel, chain_grid(n), dim3(256), 0, s, A);
  return hipGetLastError();
}
hipError_t launch_da_propose(const DaArgs* A, uint32_t n, uint32_t t,
                             hipStream_t s) {
  hipLaunchKernelGGL(da_propose_kernel,
                     dim3((n + kProposeWaves - 1) / kProposeWaves), dim3(256), 0,
                     s, A, t);
  return hipGetLastError();
}
hipError_t launch_da_stage1(const DaArgs* A, uint32_t t, hipStream_t s) {
  hipLaunchKernelGGL(da_stage1_kernel, dim3(1), dim3(1024), 0, s, A, t);
  return hipGetLastError();
}
hipError_t launch_da_draws(const float* draws_it, float* draws, uint32_t n_iter,
                           uint32_t rows, hipStream_t s) {
  if (n_iter == 0 || rows == 0) return hipSuccess;
  hipLaunchKernelGGL(da_draws_kernel, dim3((rows + 31) / 32, (n_iter + 31) / 32),
                     dim3(256), 0, s, draws_it, draws, n_iter, rows);
  return hipGetLastError();
}
hipError_t launch_da_stage2(const DaArgs* A, uint32_t n, uint32_t t,
                            hipStream_t s) {
  hipLaunchKernelGGL(da_stage2_kernel, chain_grid(n), dim3(256), 0, s, A, t);
  return hipGetLastError();
}
hipError_t launch_da_features(const double* points, uint32_t n, uint32_t n_in,
                              const double* mean, const double* w,
                              uint16_t* Xb, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(da_features_kernel, chain_grid(n), dim3(256), 0, s, points,
                     n, n_in, mean, w, Xb);
  return hipGetLastError();
}

}  // namespace gpum
