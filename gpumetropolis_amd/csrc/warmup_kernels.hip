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

// One thread per chain.
__global__ void warmup_chains_kernel(WarmupDev w, uint32_t b, uint32_t T) {
#pragma clang fp contract(off)
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= w.n_chains || w.ctl[kWarmFrozenAt] != 0u) return;
  const uint32_t np = w.np, C = w.n_chains;
  auto x_at = [&](uint32_t t, uint32_t j) {
    return (double)w.draws[t + (size_t)T * (c + (size_t)C * j)];
  };
  double* mean = w.w_mean + (size_t)c * np;
  double* m2 = w.w_m2 + (size_t)c * np;
  double* bm = w.b_mean + (size_t)c * np;
  double* bq = w.b_m2 + (size_t)c * np * np;
  // The chain's Welford moments absorb its draws one by one
  // (R/adaptive.R:31-55); the diagonal is all the scale reads.
  double n = w.w_n[c];
  for (uint32_t t = 0; t < T; ++t) {
    n += 1.0;
    for (uint32_t j = 0; j < np; ++j) {
      const double x = x_at(t, j);
      const double d1 = x - mean[j];
      mean[j] = mean[j] + d1 / n;
      m2[j] = m2[j] + d1 * (x - mean[j]);
    }
  }
  w.w_n[c] = n;
  // The batch's mean and deviation products, in draw order, for the
  // pooled accumulator.
  for (uint32_t j = 0; j < np; ++j) {
    double s = x_at(0, j);
    for (uint32_t t = 1; t < T; ++t) s = s + x_at(t, j);
    bm[j] = s / (double)T;
  }
  for (uint32_t i = 0; i < np; ++i)
    for (uint32_t j = 0; j < np; ++j) bq[i * np + j] = 0.0;
  for (uint32_t t = 0; t < T; ++t)
    for (uint32_t i = 0; i < np; ++i) {
      const double di = x_at(t, i) - bm[i];
      for (uint32_t j = 0; j < np; ++j)
        bq[i * np + j] = bq[i * np + j] + di * (x_at(t, j) - bm[j]);
    }
  // Robbins-Monro on the batch acceptance (R/adaptive.R:78-85); the rate
  // crosses the reference's FFI as f32.
  const double acc = (double)w.accept[c];
  const double g = w.gain[b - 1];
  const double sc = exp(log(w.scale[c]) + g * (acc - w.target));
  w.scale[c] = sc;
  // Next batch: the last draw, and scale * sqrt(diag cov) per dimension
  // (the seed covariance until two draws).
  for (uint32_t j = 0; j < np; ++j) {
    w.init[(size_t)c * np + j] = x_at(T - 1, j);
    const double var = n >= 2.0 ? m2[j] / (n - 1.0) : w.w_sig[(size_t)c * np + j];
    w.sd[(size_t)c * np + j] = sc * sqrt(fmax(var, 1e-12));
  }
  w.acc_hist[(size_t)(b - 1) * C + c] = w.accept[c];
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

// One workgroup of 256 threads; thread (i, j) < np x np owns element
// (i, j) of the pooled M2 and follows means i and j through the merge.
__global__ void warmup_pool_kernel(WarmupDev w, uint32_t b, uint32_t T) {
#pragma clang fp contract(off)
  if (w.ctl[kWarmFrozenAt] != 0u) return;
  const uint32_t np = w.np, C = w.n_chains;
  const uint32_t tid = threadIdx.x;
  __shared__ double cov[16 * 16];
  __shared__ double n_sh;
  const double n0 = w.p_n[0];
  double m2 = 0.0, mi = 0.0, mj = 0.0;
  uint32_t i = 0, j = 0;
  const bool mine = tid < np * np;
  if (mine) {
    i = tid / np;
    j = tid - i * np;
    m2 = w.p_m2[tid];
    mi = w.p_mean[i];
    mj = w.p_mean[j];
  }
  __syncthreads();  // every old value is read before any is written
  if (mine) {
    // Each chain's batch merged in chain order: the parallel-combination
    // identity of Chan, Golub and LeVeque, equal in exact arithmetic to
    // folding the draws in one by one (R/adaptive.R:170-177).
    double n = n0;
    const double Td = (double)T;
    for (uint32_t c = 0; c < C; ++c) {
      const double n_new = n + Td;
      const double di = w.b_mean[(size_t)c * np + i] - mi;
      const double dj = w.b_mean[(size_t)c * np + j] - mj;
      const double f1 = Td / n_new;
      mi = mi + di * f1;
      mj = mj + dj * f1;
      const double f2 = (n * Td) / n_new;
      m2 = (m2 + w.b_m2[(size_t)c * np * np + tid]) + (di * dj) * f2;
      n = n_new;
    }
    w.p_m2[tid] = m2;
    if (i == j) w.p_mean[i] = mi;
    if (tid == 0) n_sh = n;
    cov[tid] = m2;
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
  hipLaunchKernelGGL(warmup_chains_kernel, chain_grid(w.n_chains), dim3(256),
                     0, s, w, b, batch_len);
  return hipGetLastError();
}

hipError_t launch_warmup_pool(const WarmupDev& w, uint32_t b,
                              uint32_t batch_len, hipStream_t s) {
  hipLaunchKernelGGL(warmup_pool_kernel, dim3(1), dim3(256), 0, s, w, b,
                     batch_len);
  return hipGetLastError();
}

hipError_t launch_warmup_post(const WarmupDev& w, uint32_t b, hipStream_t s) {
  hipLaunchKernelGGL(warmup_post_kernel, chain_grid(w.n_chains), dim3(256), 0,
                     s, w, b);
  return hipGetLastError();
}

}  // namespace gpum
