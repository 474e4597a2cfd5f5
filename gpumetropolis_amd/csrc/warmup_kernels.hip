// warmup_kernels.hip -- the between-batch step of the adaptive warmup on
// the device (R/adaptive.R:101-254; primitives R/adaptive.R:15-85 and
// 548-558).  Each operation is rounded separately, as the host
// orchestrator computes it.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "warmup.h"

namespace gpum {

__global__ void warmup_init_kernel(WarmupDev w, const double* __restrict__ sd0) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t np = w.np;
  if (c < w.n_chains) {
    w.w_n[c] = 0.0;
    for (uint32_t j = 0; j < np; ++j) {
      const double s = sd0[(size_t)c * np + j];
      w.w_mean[(size_t)c * np + j] = 0.0;
      w.w_m2[(size_t)c * np + j] = 0.0;
      w.w_sig[(size_t)c * np + j] = s * s;
      w.sd[(size_t)c * np + j] = s;
    }
    w.scale[c] = w.scale0;
    w.prev_scale[c] = w.scale0;
    for (uint32_t k = 0; k < np * np; ++k) w.L[(size_t)c * np * np + k] = 0.0;
  }
  if (c == 0) {
    w.p_n[0] = 0.0;
    for (uint32_t j = 0; j < np; ++j) {
      w.p_mean[j] = 0.0;
      w.p_sig[j] = sd0[j] * sd0[j];
      for (uint32_t k = 0; k < np; ++k) {
        w.p_m2[j * np + k] = 0.0;
        w.L_pool[j * np + k] = 0.0;
      }
    }
    for (uint32_t k = 0; k < kWarmCtlWords; ++k) w.ctl[k] = 0u;
  }
}

// Chains per workgroup of warmup_chains_kernel: whole chains of np x np
// threads each (np <= 16).
static __host__ __device__ inline uint32_t warm_chains_per_block(uint32_t np) {
  return 256u / (np * np);
}

// One thread per element (i, j) of each chain's np x np batch M2, the
// threads of a chain adjacent in one workgroup: thread (c, i, j) sums the
// batch's deviation products of dimensions i and j, and the threads
// (c, j, j) also carry dimension j's Welford moments.  Every running value
// lives in registers.
__global__ void warmup_chains_kernel(WarmupDev w, uint32_t b, uint32_t T) {
#pragma clang fp contract(off)
  if (w.ctl[kWarmFrozenAt] != 0u) return;  // uniform over the grid
  const uint32_t np = w.np, nn = np * np, C = w.n_chains;
  const uint32_t c = blockIdx.x * warm_chains_per_block(np) + threadIdx.x / nn;
  const uint32_t e = threadIdx.x % nn;
  const uint32_t i = e / np, j = e % np;
  const bool live = c < C;
  double n = 0.0, sc = 0.0;
  if (live) {
    const float* xi = w.draws + (size_t)T * (c + (size_t)C * i);
    const float* xj = w.draws + (size_t)T * (c + (size_t)C * j);
    // The batch's means and deviation products, in draw order, for the
    // pooled accumulator.
    double si = (double)xi[0], sj = (double)xj[0];
    for (uint32_t t = 1; t < T; ++t) {
      si = si + (double)xi[t];
      sj = sj + (double)xj[t];
    }
    const double bmi = si / (double)T, bmj = sj / (double)T;
    double q = 0.0;
    for (uint32_t t = 0; t < T; ++t) {
      const double di = (double)xi[t] - bmi;
      q = q + di * ((double)xj[t] - bmj);
    }
    w.b_m2[(size_t)c * nn + e] = q;
    n = w.w_n[c];
    // Robbins-Monro on the batch acceptance (R/adaptive.R:78-85); the rate
    // crosses the reference's FFI as f32.
    const double acc = (double)w.accept[c];
    const double g = w.gain[b - 1];
    sc = exp(log(w.scale[c]) + g * (acc - w.target));
    if (i == j) {
      w.b_mean[(size_t)c * np + j] = bmj;
      // Dimension j's Welford moments absorb the draws one by one
      // (R/adaptive.R:31-55); the diagonal is all the scale reads.
      double mean = w.w_mean[(size_t)c * np + j];
      double m2 = w.w_m2[(size_t)c * np + j];
      for (uint32_t t = 0; t < T; ++t) {
        n += 1.0;
        const double x = (double)xj[t];
        const double d1 = x - mean;
        mean = mean + d1 / n;
        m2 = m2 + d1 * (x - mean);
      }
      w.w_mean[(size_t)c * np + j] = mean;
      w.w_m2[(size_t)c * np + j] = m2;
      // Next batch: the last draw, and scale * sqrt(diag cov) per
      // dimension (the seed covariance until two draws).
      w.init[(size_t)c * np + j] = (double)xj[T - 1];
      const double var =
          n >= 2.0 ? m2 / (n - 1.0) : w.w_sig[(size_t)c * np + j];
      w.sd[(size_t)c * np + j] = sc * sqrt(fmax(var, 1e-12));
    }
  }
  // Every thread of the chain has read its count and scale before thread
  // (c, 0, 0) replaces them.
  __syncthreads();
  if (live && e == 0) {
    w.w_n[c] = n;
    w.scale[c] = sc;
    w.acc_hist[(size_t)(b - 1) * C + c] = w.accept[c];
  }
}

// Lower-triangular L with L L' = S + reg I, reg = 1e-8 * mean(diag S)
// (R/adaptive.R:548-558); the diagonal sqrt(max(diag, reg)) when a pivot
// is not positive.
__device__ void warmup_cholesky(const double* S, uint32_t np, double* L) {
#pragma clang fp contract(off)
  double dsum = 0.0;
  for (uint32_t j = 0; j < np; ++j) dsum = dsum + S[j * np + j];
  double reg = 1e-8 * (dsum / (double)np);
  if (!isfinite(reg) || reg <= 0.0) reg = 1e-8;
  bool ok = true;
  for (uint32_t k = 0; k < np * np; ++k) L[k] = 0.0;
  for (uint32_t j = 0; j < np && ok; ++j) {
    double s = S[j * np + j] + reg;
    for (uint32_t k = 0; k < j; ++k) s = s - L[j * np + k] * L[j * np + k];
    if (!(s > 0.0)) {
      ok = false;
      break;
    }
    const double ljj = sqrt(s);
    L[j * np + j] = ljj;
    for (uint32_t i = j + 1; i < np; ++i) {
      double a = S[i * np + j];
      for (uint32_t k = 0; k < j; ++k) a = a - L[i * np + k] * L[j * np + k];
      L[i * np + j] = a / ljj;
    }
  }
  if (!ok) {
    for (uint32_t k = 0; k < np * np; ++k) L[k] = 0.0;
    for (uint32_t j = 0; j < np; ++j)
      L[j * np + j] = sqrt(fmax(S[j * np + j] + reg, reg));
  }
}

// Groups of consecutive chains for the pooled merge: at most kPoolGroups,
// and as many as kPoolThreads threads hold with np x np threads a group.
constexpr uint32_t kPoolThreads = 1024;
constexpr uint32_t kPoolGroups = 64;

// One workgroup of kPoolThreads threads.  The batch moments of the chains
// are merged into the pooled accumulator by the parallel-combination
// identity of Chan, Golub and LeVeque, equal in exact arithmetic to folding
// the draws in one by one (R/adaptive.R:170-177), in two levels of fixed
// order: G groups of S consecutive chains each merge their chains in chain
// order, np x np threads a group (thread (i, j) owns element (i, j) of the
// group's M2 and follows means i and j), and then the np x np threads of
// group 0 merge the G group moments in group order into the pooled ones.
// Every count is an integer held exactly in FP64.
__global__ void __launch_bounds__(kPoolThreads)
warmup_pool_kernel(WarmupDev w, uint32_t b, uint32_t T) {
#pragma clang fp contract(off)
  if (w.ctl[kWarmFrozenAt] != 0u) return;
  const uint32_t np = w.np, nn = np * np, C = w.n_chains;
  const uint32_t tid = threadIdx.x;
  const uint32_t G = min(min(kPoolGroups, kPoolThreads / nn), C);
  const uint32_t S = (C + G - 1) / G;  // chains per group (the last fewer)
  __shared__ double cov[16 * 16];
  __shared__ double n_sh;
  // Group moments: G x np means and G x nn M2, each at most kPoolThreads
  // entries since G nn <= kPoolThreads and np <= nn.
  __shared__ double g_mean[kPoolThreads], g_m2[kPoolThreads];
  // Merge factors of the k-th chain of a group: T / ((k + 1) T) and
  // k T T / ((k + 1) T), computed once for every group.
  __shared__ double f1s[kPoolThreads], f2s[kPoolThreads];
  const double Td = (double)T;
  const uint32_t g = tid / nn, e = tid - g * nn;
  const uint32_t i = e / np, j = e - (e / np) * np;
  for (uint32_t base = 0; base < S; base += kPoolThreads) {
    const uint32_t cnt = min(kPoolThreads, S - base);
    __syncthreads();  // the previous chunk's factors are no longer read
    if (tid < cnt) {
      const double n = (double)(base + tid) * Td;
      const double n_new = n + Td;
      f1s[tid] = Td / n_new;
      f2s[tid] = (n * Td) / n_new;
    }
    __syncthreads();
    if (g < G) {
      double mi = 0.0, mj = 0.0, m2 = 0.0;
      if (base != 0) {
        mi = g_mean[g * np + i];
        mj = g_mean[g * np + j];
        m2 = g_m2[g * nn + e];
      }
      const uint32_t c0 = g * S + base;
      const uint32_t c1 = min(c0 + cnt, min((g + 1) * S, C));
      for (uint32_t c = c0; c < c1; ++c) {
        const uint32_t k = c - c0;
        const double di = w.b_mean[(size_t)c * np + i] - mi;
        const double dj = w.b_mean[(size_t)c * np + j] - mj;
        mi = mi + di * f1s[k];
        mj = mj + dj * f1s[k];
        m2 = (m2 + w.b_m2[(size_t)c * nn + e]) + (di * dj) * f2s[k];
      }
      // Each thread writes only the entries it alone owns; the reads
      // above of the same entries are by the same thread.
      if (i == j) g_mean[g * np + i] = mi;
      g_m2[g * nn + e] = m2;
    }
  }
  __syncthreads();
  if (tid < nn) {
    const double n0 = w.p_n[0];
    double mi = w.p_mean[i], mj = w.p_mean[j], m2 = w.p_m2[e];
    for (uint32_t h = 0; h < G; ++h) {
      const uint32_t c0 = h * S;
      if (c0 >= C) break;
      const double nb = (double)(min(c0 + S, C) - c0) * Td;
      const double n = n0 + (double)c0 * Td;
      const double n_new = n + nb;
      const double di = g_mean[h * np + i] - mi;
      const double dj = g_mean[h * np + j] - mj;
      const double f1 = nb / n_new;
      mi = mi + di * f1;
      mj = mj + dj * f1;
      m2 = (m2 + g_m2[h * nn + e]) + (di * dj) * ((n * nb) / n_new);
    }
    w.p_m2[e] = m2;
    if (i == j) w.p_mean[i] = mi;
    if (tid == 0) n_sh = n0 + (double)C * Td;
    cov[e] = m2;
  }
  __syncthreads();
  if (tid != 0) return;
  double n = n_sh;
  w.p_n[0] = n;
  // .am_covariance: M2 / (n - 1), the seed until two draws.
  for (uint32_t k = 0; k < np * np; ++k) {
    const uint32_t r = k / np, q = k - r * np;
    cov[k] = n >= 2.0 ? cov[k] / (n - 1.0) : (r == q ? w.p_sig[r] : 0.0);
  }
  if (b == w.reset_batch) {
    // Mid-warmup restart: the pooled accumulator restarts from the pooled
    // variances (the per-chain ones follow in warmup_post_kernel).
    for (uint32_t r = 0; r < np; ++r) {
      const double v = fmax(cov[r * np + r], 1e-12);
      w.diag_now[r] = v;
      w.p_sig[r] = v;
      w.p_mean[r] = 0.0;
      for (uint32_t q = 0; q < np; ++q) w.p_m2[r * np + q] = 0.0;
    }
    n = 0.0;
    w.p_n[0] = 0.0;
  }
  if (np > 1 && n >= 2.0) {
    warmup_cholesky(cov, np, w.L_pool);
    w.ctl[kWarmHasL] = 1u;
  }
  w.ctl[kWarmBatches] = b;
  w.ctl[kWarmUsed] += w.sizes[b - 1];
  // Early freeze (R/adaptive.R:224-239): acceptance near the target and
  // scalars that stopped moving, two batches in a row.
  if (w.auto_stop) {
    const uint32_t from = w.reset_batch + 5 > 10 ? w.reset_batch + 5 : 10;
    if (b >= from) {
      double s = 0.0, drift = 0.0;
      for (uint32_t c = 0; c < C; ++c) {
        s = s + fabs((double)w.accept[c] - w.target);
        drift = fmax(drift, fabs(w.scale[c] / w.prev_scale[c] - 1.0));
      }
      const bool calm = s / (double)C < 0.05 && drift < 0.03;
      w.ctl[kWarmStreak] = calm ? w.ctl[kWarmStreak] + 1u : 0u;
      if (w.ctl[kWarmStreak] >= 2u) w.ctl[kWarmFrozenAt] = b;
    }
  }
}

// One thread per chain.
__global__ void warmup_post_kernel(WarmupDev w, uint32_t b) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t fa = w.ctl[kWarmFrozenAt];
  if (c >= w.n_chains || (fa != 0u && fa < b)) return;
  const uint32_t np = w.np;
  if (b == w.reset_batch) {
    w.w_n[c] = 0.0;
    for (uint32_t j = 0; j < np; ++j) {
      w.w_mean[(size_t)c * np + j] = 0.0;
      w.w_m2[(size_t)c * np + j] = 0.0;
      w.w_sig[(size_t)c * np + j] = w.diag_now[j];
    }
    w.scale[c] = w.scale0;
  }
  const double sc = w.scale[c];
  if (np > 1 && w.ctl[kWarmHasL] != 0u)
    for (uint32_t k = 0; k < np * np; ++k)
      w.L[(size_t)c * np * np + k] = sc * w.L_pool[k];
  w.prev_scale[c] = sc;
}

static dim3 chain_grid(uint32_t n) { return dim3((n + 255) / 256); }

hipError_t launch_warmup_init(const WarmupDev& w, const double* sd0,
                              hipStream_t s) {
  hipLaunchKernelGGL(warmup_init_kernel, chain_grid(w.n_chains), dim3(256), 0,
                     s, w, sd0);
  return hipGetLastError();
}

hipError_t launch_warmup_chains(const WarmupDev& w, uint32_t b,
                                uint32_t batch_len, hipStream_t s) {
  const uint32_t k = warm_chains_per_block(w.np);
  hipLaunchKernelGGL(warmup_chains_kernel, dim3((w.n_chains + k - 1) / k),
                     dim3(k * w.np * w.np), 0, s, w, b, batch_len);
  return hipGetLastError();
}

hipError_t launch_warmup_pool(const WarmupDev& w, uint32_t b,
                              uint32_t batch_len, hipStream_t s) {
  hipLaunchKernelGGL(warmup_pool_kernel, dim3(1), dim3(kPoolThreads), 0, s, w, b,
                     batch_len);
  return hipGetLastError();
}

hipError_t launch_warmup_post(const WarmupDev& w, uint32_t b, hipStream_t s) {
  hipLaunchKernelGGL(warmup_post_kernel, chain_grid(w.n_chains), dim3(256), 0,
                     s, w, b);
  return hipGetLastError();
}

}  // namespace gpum
