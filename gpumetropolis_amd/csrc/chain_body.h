// chain_body.h -- device code of the chain, points and batched kernels,
// generic over the per-observation evaluator E:
//   E::kOpt         observations per lane per chunk;
//   E::kCw          chains per wave: one wave evaluates kCw chains on every
//                   observation it reads from LDS (chain j's uniform table
//                   is U + j * H.n_uniform_pad);
//   E::kPairs       whether the evaluator sums two chunks per FP32 partial
//                   when H.unit_chunks is 2 (programs without spill slots);
//   E::kUnit        chunks per FP32 partial fixed at compile time (1 or 2;
//                   0: H.unit_chunks);
//   E::kScalarCopy  tile copies addressed on the scalar unit (buffer
//                   descriptors) rather than with per-lane 64-bit addresses
//                   (the interpreters, short of scalar registers, keep the
//                   latter);
//   E::acc_state(H, U)
//                   what a data pass keeps in registers from the uniform
//                   table (e.g. the uniforms a compiled program reads), read
//                   once per pass and held across its tiles;
//   E::run(H, state, U, tile, ld, lane, T)
//                   evaluate the program on one chunk into T[kCw][kOpt]
//                   (`tile`: the chunk, `ld`: its column stride, 64*kOpt);
//   E::mode(H)      how a chunk is summed (SumMode);
//   E::run_acc(H, state, U, tile, ld, lane, s)
//                   SUM_AFFINE: the chunk's products added straight to the
//                   sums s[kCw] (the caller starts them);
//   E::finish(H, U, s, with_const)
//                   apply the sum tail to one chain's data-pass sum;
//   E::Params::hoist(code, n, consts, p, U), E::Params::prior(code, n,
//                   consts, p)
//                   the FP64 parameter programs (hoisted values into the
//                   uniform table; the prior).
// The interpreters (VmEval, and VmAffineEval for affine programs) and code
// generated per model share this file.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "chain_kernel.h"
#include "device_vm.h"
#include "math64.h"
#include "philox.h"
#include "step_common.h"

namespace gpum {


// Per-chain record in LDS (doubles):
//   state[np] prop[np] z[np] sd[np] wmean[np] wm2[np]
//   cur prior_lp accepts wn scale batch_accepts
struct Rec {
  double* state;
  double* prop;
  double* z;
  double* sd;
  double* wmean;
  double* wm2;
  double* tail;  // cur, prior_lp, accepts, wn, scale, batch_accepts
  __device__ double& cur() { return tail[0]; }
  __device__ double& prior_lp() { return tail[1]; }
  __device__ double& accepts() { return tail[2]; }
  __device__ double& wn() { return tail[3]; }
  __device__ double& scale() { return tail[4]; }
  __device__ double& batch_acc() { return tail[5]; }
};

__device__ __forceinline__ Rec make_rec(double* st, int np) {
  Rec r;
  r.state = st;
  r.prop = st + np;
  r.z = st + 2 * np;
  r.sd = st + 3 * np;
  r.wmean = st + 4 * np;
  r.wm2 = st + 5 * np;
  r.tail = st + 6 * np;
  return r;
}

// Batch structure of a run split into batches of `every` iterations, the
// last absorbing the remainder (R/adaptive.R:284-289).  Returns the 1-based
// batch that ends after iteration t, or 0.
__device__ __forceinline__ uint32_t batch_ending_at(uint32_t t, uint32_t n_iter,
                                                    uint32_t every,
                                                    uint32_t n_batches) {
  if (every == 0) return 0;
  if (t + 1 == n_iter) return n_batches;
  if ((t + 1) % every == 0 && (t + 1) / every < n_batches) return (t + 1) / every;
  return 0;
}

// Finish iteration it-1 with the data term `ll` of the proposal
// (cpu_native.rs:785-799): tempered Metropolis decision in FP64, record the
// f32 draw; during adaptation batches, fold the draw into the chain's
// Welford accumulator (R/adaptive.R:31-42, diagonal part).
__device__ __noinline__ void chain_finish_iter(const ChainArgs* __restrict__ A,
                                               uint32_t it, uint32_t chain,
                                               double* st, double ll) {
  const int np = (int)A->n_params;
  const uint32_t C = A->n_chains;
  const uint32_t T_iter = A->n_iter;
  const uint32_t spi = A->mode == 1 ? 2u * np + 4u : 2u * np + 1u;
  Rec r = make_rec(st, np);
  double plp = ll + r.prior_lp();
  if (it == 0) {
    r.cur() = plp;
    if (A->init_log_post) A->init_log_post[chain] = plp;
    return;
  }
  const uint32_t t = it - 1;
  StepUniforms us{A->seed, A->stream_id, A->iter_offset + t, chain};
  double u = us.get(spi - 1);
  bool acc_ok = log(u) < (plp - r.cur()) / A->temperatures[chain];
  if (A->trace_plp) A->trace_plp[(size_t)t * C + chain] = plp;
  if (A->trace_ll) A->trace_ll[(size_t)t * C + chain] = ll;
  if (A->trace_prop)
    for (int j = 0; j < np; ++j)
      A->trace_prop[((size_t)t * C + chain) * np + j] = r.prop[j];
  if (acc_ok) {
    for (int j = 0; j < np; ++j) r.state[j] = r.prop[j];
    r.cur() = plp;
    r.accepts() += 1.0;
    r.batch_acc() += 1.0;
  }
  if (A->trace_accept) A->trace_accept[(size_t)t * C + chain] = acc_ok ? 1 : 0;
  const size_t ld = A->draws_ld ? A->draws_ld : T_iter;
  for (int j = 0; j < np; ++j)
    A->draws[t + ld * (chain + (size_t)C * j)] = (float)r.state[j];
  if (A->adapt_batches > 0 && A->swap_every > 0) {
    const uint32_t b = min(t / A->swap_every + 1, A->n_batches);
    if (b <= A->adapt_batches) {
      // Welford over the recorded f32 draws, as the reference's host sees
      // them (R/adaptive.R:31-42).
      r.wn() += 1.0;
      for (int j = 0; j < np; ++j) {
        double x = (double)(float)r.state[j];
        double d1 = x - r.wmean[j];
        r.wmean[j] += d1 / r.wn();
        r.wm2[j] += d1 * (x - r.wmean[j]);
      }
    }
  }
}

// End of batch b for one chain: Robbins-Monro on the batch acceptance and
// the per-dimension scale sqrt(diag(cov)) (R/adaptive.R:342-356), while b is
// an adaptation batch; then reset the batch counter.
__device__ __noinline__ void chain_adapt(const ChainArgs* __restrict__ A,
                                         uint32_t b, uint32_t bsize,
                                         uint32_t chain, double* st) {
  const int np = (int)A->n_params;
  Rec r = make_rec(st, np);
  // accept_rate crosses the reference's FFI as f32 (chain_kernel.rs:353).
  const double acc = (double)((float)r.batch_acc() / (float)bsize);
  if (A->acc_hist) A->acc_hist[(size_t)(b - 1) * A->n_chains + chain] = (float)acc;
  r.batch_acc() = 0.0;
  if (b > A->adapt_batches) return;
  const double g = pow((double)(A->batch0 + b), -2.0 / 3.0);
  r.scale() = exp(log(r.scale()) + g * (acc - target_accept(np)));
  for (int j = 0; j < np; ++j) {
    // .am_covariance: the seed Sigma (initial sd^2) until two draws.
    double var = r.wn() >= 2.0 ? r.wm2[j] / (r.wn() - 1.0)
                               : A->proposal_sd[(size_t)chain * np + j] *
                                     A->proposal_sd[(size_t)chain * np + j];
    r.sd[j] = r.scale() * sqrt(fmax(var, 1e-12));
  }
}

// Replica-exchange sweep of batch b over the ladder groups of this
// workgroup (R/adaptive.R:362-391): odd batches try pairs (0,1), (2,3), ...,
// even batches (1,2), (3,4), ...; a pair swaps states and log-posteriors
// when log u < (lp[i+1] - lp[i]) (1/T_i - 1/T_{i+1}).  The uniform for the
// k-th pair of the sweep of group g is word 0 of Philox block
// (b, k, g, TAG_SWAP).
__device__ __noinline__ void pt_sweep(const ChainArgs* __restrict__ A,
                                      uint32_t b, uint32_t chain0, int n_local,
                                      double* st_all, int SW) {
  const int np = (int)A->n_params;
  const uint32_t G = A->swap_group;
  const uint32_t bg = A->batch0 + b;  // global batch index
  for (int g0 = 0; g0 + (int)G <= n_local; g0 += (int)G) {
    const uint32_t gchain0 = chain0 + (uint32_t)g0;
    if (gchain0 + G > A->n_chains) break;
    const uint32_t group = gchain0 / G;
    const int pairs = (int)G - 1;
    if (pairs < 1) continue;
    int start = (bg % 2u == 1u) ? 0 : (pairs >= 2 ? 1 : 0);
    int k = 0;
    for (int i = start; i < pairs; i += 2, ++k) {
      U4 w = philox4x32_10(U4{bg, (uint32_t)k, group, TAG_SWAP}, A->seed,
                           A->stream_id);
      double u = u32_to_unit(w.x);
      Rec lo = make_rec(st_all + (g0 + i) * SW, np);
      Rec hi = make_rec(st_all + (g0 + i + 1) * SW, np);
      const double t_lo = A->temperatures[gchain0 + i];
      const double t_hi = A->temperatures[gchain0 + i + 1];
      double delta = (hi.cur() - lo.cur()) * (1.0 / t_lo - 1.0 / t_hi);
      bool ok = isfinite(delta) && log(u) < delta;
      if (ok) {
        for (int j = 0; j < np; ++j) {
          double x = lo.state[j];
          lo.state[j] = hi.state[j];
          hi.state[j] = x;
        }
        double c = lo.cur();
        lo.cur() = hi.cur();
        hi.cur() = c;
      }
      if (A->swap_hist)
        A->swap_hist[(size_t)(b - 1) * A->n_chains + gchain0 + i] = ok ? 1 : 0;
    }
  }
}

// The proposal's parameter programs: the hoisted program into the uniform
// table and the prior (the evaluator's own FP64 code).
template <class E>
__device__ __forceinline__ void step_params(const ChainArgs* __restrict__ A,
                                            double* st, float* U) {
  Rec r = make_rec(st, (int)A->n_params);
  E::hoist(A->hoist, A->n_hoist, A->consts, r.prop, U);
  r.prior_lp() = E::prior(A->prior, A->n_prior, A->prior_consts, r.prop);
}

// The proposal's normal deviates of iteration it, in parallel: lane
// 16 k + j of a chain group's first wave draws z_j of the group's chain k
// (Box-Muller on the Philox slots of cpu_native.rs:728-735, after the pair
// and the scale in mode 1) into the chain's record, where
// chain_prepare_iter reads them.  n_params <= 16 and at most four chains
// per wave, so every (chain, parameter) has its lane.
__device__ __noinline__ void chain_normals(const ChainArgs* __restrict__ A,
                                           uint32_t it, uint32_t chain_first,
                                           int n_group, int lane,
                                           double* st_first, int SW) {
  const int np = (int)A->n_params;
  const int k = lane >> 4, j = lane & 15;
  const uint32_t chain = chain_first + (uint32_t)k;
  if (k >= n_group || j >= np || chain >= A->n_chains) return;
  const uint32_t zo = A->mode == 1 ? 3u : 0u;
  const double two_pi = 6.283185307179586;
  StepUniforms us{A->seed, A->stream_id, A->iter_offset + it, chain};
  const double u1 = us.get(zo + 2 * j), u2 = us.get(zo + 2 * j + 1);
  make_rec(st_first + k * SW, np).z[j] = sqrt(-2.0 * log(u1)) * cos(two_pi * u2);
}

// Prepare iteration it: the proposal (cpu_native.rs:726-784) from the
// deviates chain_normals drew, then its parameter programs.  One call per
// iteration from the kernel body (each call saves the body's live
// registers).
template <class E>
__device__ __noinline__ void chain_prepare_iter(const ChainArgs* __restrict__ A,
                                                uint32_t it, uint32_t chain,
                                                double* st, float* U) {
  const int np = (int)A->n_params;
  const uint32_t C = A->n_chains;
  const uint32_t mode = A->mode;
  Rec r = make_rec(st, np);
  StepUniforms us{A->seed, A->stream_id, A->iter_offset + it, chain};
  const double* sd = r.sd;
  if (mode == 0) {
    for (int j = 0; j < np; ++j) r.prop[j] = r.state[j] + sd[j] * r.z[j];
  } else if (mode == 2) {
    const double* L = A->proposal_l + (size_t)chain * np * np;
    for (int k = 0; k < np; ++k) {
      double a = r.state[k];
      for (int j = 0; j <= k; ++j) a += L[k * np + j] * r.z[j];
      r.prop[k] = a;
    }
  } else {
    // DE path A: the pair comes from the frozen batch-start population in
    // `init`; with probability 0.1 the scale is 1.
    uint32_t ia = (uint32_t)(us.get(0) * (double)C);
    if (ia >= C) ia = C - 1;
    uint32_t ib = (uint32_t)(us.get(1) * (double)C);
    if (ib >= C) ib = C - 1;
    if (ib == ia) ib = (ib + 1) % C;
    double g = us.get(2) < 0.1 ? 1.0 : A->gamma;
    for (int j = 0; j < np; ++j) {
      double diff = A->init[(size_t)ia * np + j] - A->init[(size_t)ib * np + j];
      r.prop[j] = r.state[j] + g * diff + A->de_noise * sd[j] * r.z[j];
    }
  }
  step_params<E>(A, st, U);
}

template <class E>
__device__ __noinline__ void chain_init(const ChainArgs* __restrict__ A,
                                        uint32_t chain, double* st, float* U) {
  const int np = (int)A->n_params;
  Rec r = make_rec(st, np);
  for (int j = 0; j < np; ++j) {
    double v = A->init[(size_t)chain * np + j];
    r.state[j] = v;
    r.prop[j] = v;
    r.sd[j] = A->proposal_sd[(size_t)chain * np + j];
    r.wmean[j] = 0.0;
    r.wm2[j] = 0.0;
  }
  r.accepts() = 0.0;
  r.wn() = 0.0;
  r.scale() = A->init_scale ? A->init_scale[chain] : 1.0;
  r.batch_acc() = 0.0;
  step_params<E>(A, st, U);
}

__device__ __noinline__ void chain_finish(const ChainArgs* __restrict__ A,
                                          uint32_t chain, double* st) {
  const int np = (int)A->n_params;
  Rec r = make_rec(st, np);
  const uint32_t T_iter = A->n_iter;
  A->accept_rate[chain] =
      T_iter == 0 ? 0.0f : (float)r.accepts() / (float)T_iter;
  A->last_log_post[chain] = r.cur();
  if (A->final_state)
    for (int j = 0; j < np; ++j) A->final_state[(size_t)chain * np + j] = r.state[j];
  if (A->final_sd)
    for (int j = 0; j < np; ++j) A->final_sd[(size_t)chain * np + j] = r.sd[j];
  if (A->final_scale) A->final_scale[chain] = r.scale();
}

// Streamed tiles live in two static LDS buffers.  Being distinct LDS
// objects, the compiler can tell that the waves' reads of one buffer do not
// depend on the asynchronous copy into the other, so the copy overlaps the
// evaluation instead of being waited for at the first LDS read.
__shared__ float4 g_tile_a[kTileF4];
__shared__ float4 g_tile_b[kTileF4];

// LDS tile layout, chunk-major: the tile's observations in chunks of
// 64*OPT, each chunk holding its n_cols columns one after another (64*OPT
// floats each, consecutive observations, as in HBM).  Within a chunk every
// column is then a constant offset from one base, so a compiled program's
// reads of its columns carry immediate offsets, and a chunk costs one
// address step however many columns it reads.  A (chunk, column) segment
// is 16*OPT float4s, a multiple of 64 (OPT is 4, 8 or 16).
// HBM float4 of LDS float4 i of tile ti (n4 float4s per column of a tile,
// ld4 float4s per column in HBM).
template <int OPT>
__device__ __forceinline__ size_t tile_src4(uint32_t i, uint32_t n_cols,
                                            uint32_t ld4, uint32_t n4,
                                            uint32_t ti) {
  constexpr uint32_t seg4 = 16u * OPT;
  const uint32_t seg = i / seg4, sub = i - seg * seg4;
  const uint32_t chunk = seg / n_cols, col = seg - chunk * n_cols;
  return (size_t)col * ld4 + (size_t)ti * n4 + (size_t)chunk * seg4 + sub;
}

// Asynchronous tile copy: a buffer load with the LDS flag moves 16 bytes
// per lane from HBM straight into LDS (no VGPR staging); lane l of wave w
// writes float4 i0 + l where i0 = k*blockDim + 64*w is wave-uniform.  A tile holds a
// multiple of 64 float4s, and 64 consecutive LDS float4s lie in one
// (chunk, column) segment, so a wave's slots are all inside the tile or all
// outside and read 64 consecutive HBM float4s.  Streaming launches use 1024
// threads, so a tile takes at most kTileIssues copies per wave.  Slots past
// the tile repeat its last 64 float4s (identical bytes to identical
// addresses), which keeps the sequence free of branches.
struct TileCopy {
  uint32_t base;            // LDS float4 index of slot 0 (64 * wave)
  uint32_t last;            // the tile's last slot (tile4 - 64)
  size_t src[kTileIssues];  // each slot's HBM float4 offset within tile 0
  // LDS float4 index of slot k, recomputed per copy on the scalar unit
  // (fewer scalar registers held across the loop).
  __device__ __forceinline__ uint32_t lds(uint32_t k) const {
    return min(base + k * 1024u, last);
  }
};

// The wave's copy slots, computed once per launch: they are wave-uniform
// (scalar registers), and a tile's copy only adds ti * n4.
template <int OPT>
__device__ __forceinline__ TileCopy tile_copy_plan(uint32_t n_cols,
                                                   uint32_t ld4, uint32_t n4,
                                                   uint32_t tile4,
                                                   uint32_t wave_base) {
  TileCopy c;
  c.base = wave_base;
  c.last = tile4 - 64u;
#pragma unroll
  for (uint32_t k = 0; k < kTileIssues; ++k)
    c.src[k] = tile_src4<OPT>(c.lds(k), n_cols, ld4, n4, 0);
  return c;
}

// Word 3 of a raw (untyped, unswizzled) buffer descriptor on gfx9-family
// GPUs, gfx950 included.
constexpr int kRawBufferWord3 = 0x00020000;

// One tile's copy.  Each copy slot reads 64 consecutive float4s (one per
// lane) through a buffer descriptor whose base is the slot's first float4
// of tile ti: the base is computed on the scalar unit, every lane's offset
// is lane * 16 bytes, so a tile costs no vector address arithmetic and no
// data set is too large for 32-bit offsets.  The descriptor's range is the
// slot's 1024 bytes.  The loop is unrolled so that every copy in flight
// keeps its own descriptor.
__device__ __forceinline__ void issue_tile_dma(const float4* __restrict__ src,
                                               const TileCopy& c, uint32_t n4,
                                               int lane, uint32_t ti,
                                               float4* dst) {
  const uint32_t voff = (uint32_t)lane * 16u;
  // The slots' LDS indices are recomputed here on the scalar unit: the
  // empty asm keeps the compiler from hoisting them out of the tile loop
  // into scalar registers it would then have to spill.
  TileCopy cc = c;
  asm volatile("" : "+s"(cc.base));
#pragma unroll
  for (uint32_t k = 0; k < kTileIssues; ++k) {
    const float4* base = src + (c.src[k] + (size_t)ti * n4);
    const __amdgpu_buffer_rsrc_t r = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float4*>(base), (short)0, 1024, kRawBufferWord3);
    __builtin_amdgcn_raw_ptr_buffer_load_lds(
        r, (__attribute__((address_space(3))) void*)(dst + cc.lds(k)), 16,
        voff, 0, 0, 0);
  }
}

// The same copy with each lane's 64-bit address in vector registers: no
// scalar registers per slot beyond the plan's.
__device__ __forceinline__ void issue_tile_dma_flat(const float4* __restrict__ src,
                                                    const TileCopy& c,
                                                    uint32_t n4, int lane,
                                                    uint32_t ti, float4* dst) {
#pragma unroll
  for (uint32_t k = 0; k < kTileIssues; ++k) {
    const float4* base = src + (c.src[k] + (size_t)ti * n4);
    __builtin_amdgcn_global_load_lds(
        (const void*)(base + lane),
        (__attribute__((address_space(3))) void*)(dst + c.lds(k)), 16, 0, 0);
  }
}

template <class E>
__device__ __forceinline__ void issue_tile_copy(const float4* __restrict__ src,
                                                const TileCopy& c, uint32_t n4,
                                                int lane, uint32_t ti,
                                                float4* dst) {
  if constexpr (E::kScalarCopy)
    issue_tile_dma(src, c, n4, lane, ti, dst);
  else
    issue_tile_dma_flat(src, c, n4, lane, ti, dst);
}

// One full chunk of 64*OPT observations added to the FP32 partials s[j] of
// the wave's kCw chains (FIRST: s starts here).  The affine form
// accumulates its products straight into s; the others evaluate the chunk
// and add its values (or their squares).
template <class E, uint32_t MODE, bool FIRST>
__device__ __forceinline__ void chunk_sum(const ChainHot& H,
                                          const typename E::AccState& ust,
                                          const float* U,
                                          const float* __restrict__ cb,
                                          uint32_t chunk, int lane,
                                          float (&s)[E::kCw]) {
  constexpr int OPT = E::kOpt, CW = E::kCw;
  if constexpr (MODE == SUM_AFFINE) {
    if constexpr (FIRST) {
#pragma unroll
      for (int j = 0; j < CW; ++j) s[j] = 0.0f;
    }
    E::run_acc(H, ust, U, cb, chunk, lane, s);
  } else {
    float T[CW][OPT];
    E::run(H, ust, U, cb, chunk, lane, T);
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      if constexpr (MODE == SUM_SQUARE) {
        if constexpr (FIRST) s[j] = 0.0f;
#pragma unroll
        for (int k = 0; k < OPT; ++k) s[j] = fmaf(T[j][k], T[j][k], s[j]);
      } else {
        if constexpr (FIRST) s[j] = T[j][0];
        else s[j] += T[j][0];
#pragma unroll
        for (int k = 1; k < OPT; ++k) s[j] += T[j][k];
      }
    }
  }
}

// A chunk of the unit that holds the end of the data, evaluated per
// observation: observations past n_obs count zero (`start`: s starts here).
template <class E, uint32_t MODE>
__device__ __forceinline__ void chunk_sum_masked(const ChainHot& H,
                                                 const typename E::AccState& ust,
                                                 const float* U,
                                                 const float* __restrict__ cb,
                                                 uint32_t chunk, uint32_t first,
                                                 bool start, int lane,
                                                 float (&s)[E::kCw]) {
  constexpr int OPT = E::kOpt, CW = E::kCw;
  float T[CW][OPT];
  E::run(H, ust, U, cb, chunk, lane, T);
#pragma unroll
  for (int j = 0; j < CW; ++j) {
    if (start) s[j] = 0.0f;
#pragma unroll
    for (int k = 0; k < OPT; ++k) {
      uint32_t obs = first + (uint32_t)(4 * lane + 256 * (k >> 2) + (k & 3));
      const float v = obs < H.n_obs ? T[j][k] : 0.0f;
      if constexpr (MODE == SUM_SQUARE)
        s[j] = fmaf(v, v, s[j]);
      else
        s[j] = s[j] + v;
    }
  }
}

// The wave's share of one tile for its kCw chains.  The tile's chunks of
// 64*OPT observations form units of `uc` consecutive chunks (E::kUnit, or
// H.unit_chunks; the tile starts on a unit boundary); the wave takes units
// part, part + wpc, ....  A unit's chunks share one FP32 partial per chain,
// added in FP64 to acc[j].  One loop per summation mode, so the
// interpreter's loop carries only its mode's registers.  `ust` holds what
// the evaluator keeps in registers for the whole data pass.
template <class E, uint32_t MODE>
__device__ __forceinline__ void tile_pass_mode(const ChainHot& H,
                                               const typename E::AccState& ust,
                                               const float* __restrict__ tile,
                                               uint32_t t0, uint32_t part,
                                               uint32_t wpc, const float* U,
                                               int lane,
                                               double (&acc)[E::kCw]) {
  constexpr int OPT = E::kOpt, CW = E::kCw;
  constexpr uint32_t chunk = 64u * OPT;
  const uint32_t nchunks = H.tile_len / chunk;
  const uint32_t cstride = H.n_cols * chunk;  // floats per chunk of the tile
  // Chunks wholly inside the data, and chunks holding any observation: all
  // of them for every tile but the last ones, so the common case costs one
  // scalar compare.  The loops below carry no per-chunk test.
  uint32_t nfull = nchunks, nlive = nchunks;
  if (t0 + H.tile_len > H.n_obs) {
    // t0 clamped to n_obs in a scalar register; the empty asm keeps the
    // compiler from folding the clamp and subtraction into a saturating
    // vector subtract.
    uint32_t t0c = min(t0, H.n_obs);
    asm volatile("" : "+s"(t0c));
    const uint32_t left = H.n_obs - t0c;
    nfull = min(nchunks, left / chunk);
    nlive = min(nchunks, (left + chunk - 1) / chunk);
  }
  // Only evaluators without spill slots pair chunks (the host sets
  // unit_chunks = 1 for the others, whose chunks are long anyway).
  const uint32_t uc =
      E::kUnit ? (uint32_t)E::kUnit : (E::kPairs ? H.unit_chunks : 1u);
  const uint32_t ufull = nfull / uc;  // units wholly inside the data
  // The unit counter is wave-uniform: kept in a scalar register, the loop
  // tests and steps it on the scalar unit.
  uint32_t u = __builtin_amdgcn_readfirstlane(part);
  if constexpr (E::kPairs) {
    const uint32_t ustride = uc * cstride;
    for (; u < ufull; u += wpc) {
      float s[CW];
      const float* cb = tile + u * ustride;
      chunk_sum<E, MODE, true>(H, ust, U, cb, chunk, lane, s);
      if (uc == 2)
        chunk_sum<E, MODE, false>(H, ust, U, cb + cstride, chunk, lane, s);
#pragma unroll
      for (int j = 0; j < CW; ++j) acc[j] += (double)s[j];
    }
  } else {
    // Unrolled by two, so one chunk's LDS reads overlap the previous
    // chunk's arithmetic.
#pragma unroll 2
    for (; u < ufull; u += wpc) {
      float s[CW];
      chunk_sum<E, MODE, true>(H, ust, U, tile + u * cstride, chunk, lane, s);
#pragma unroll
      for (int j = 0; j < CW; ++j) acc[j] += (double)s[j];
    }
  }
  // The unit holding the end of the data (one wave at most), evaluated per
  // observation with the observations past n_obs masked.  Which unit that
  // is depends on n_obs alone, so every tiling sums it alike.
  const uint32_t c0 = u * uc;
  if (u == ufull && c0 < nlive) {
    float s[CW];
    const uint32_t c1 = min(c0 + uc, nlive);
    for (uint32_t c = c0; c < c1; ++c)
      chunk_sum_masked<E, MODE>(H, ust, U, tile + c * cstride, chunk,
                                t0 + c * chunk, c == c0, lane, s);
#pragma unroll
    for (int j = 0; j < CW; ++j) acc[j] += (double)s[j];
  }
}

template <class E>
__device__ __forceinline__ void tile_pass(const ChainHot& H,
                                          const typename E::AccState& ust,
                                          const float* __restrict__ tile,
                                          uint32_t t0, uint32_t part,
                                          uint32_t wpc, const float* U,
                                          int lane, double (&acc)[E::kCw]) {
  const uint32_t mode = E::mode(H);
  if (mode == SUM_AFFINE)
    tile_pass_mode<E, SUM_AFFINE>(H, ust, tile, t0, part, wpc, U, lane, acc);
  else if (mode == SUM_SQUARE)
    tile_pass_mode<E, SUM_SQUARE>(H, ust, tile, t0, part, wpc, U, lane, acc);
  else
    tile_pass_mode<E, SUM_VALUE>(H, ust, tile, t0, part, wpc, U, lane, acc);
}

// One workgroup's chains chain0 .. chain0 + cpb - 1 through the whole run.
template <class E, bool STREAM>
// H by value: a reference to the kernel argument made the register
// allocator spill around the tile copies (a wait on the copy itself).
__device__ __forceinline__ void chain_body(const ChainArgs* __restrict__ Ap,
                                           ChainHot H, uint32_t chain0) {
  extern __shared__ __align__(16) unsigned char smem[];
  // A frozen device-side warmup: every thread reads the same word, so the
  // whole workgroup leaves before any barrier.
  if (Ap->stop && *Ap->stop != 0u) return;
  const int SW = (int)H.state_words;
  const int cpb = (int)H.cpb, wpc = (int)H.wpc;
  const int NU = (int)H.n_uniform_pad;
  double* st_all = reinterpret_cast<double*>(smem);
  double* partials = st_all + cpb * SW;
  float* Uall = reinterpret_cast<float*>(partials + cpb * wpc);

  constexpr int CW = E::kCw;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  // Wave-uniform indices: readfirstlane tells the compiler so, and the
  // chunk loops become scalar loops instead of exec-masked ones.  Wave
  // (grp, part) evaluates chains grp*CW .. grp*CW + CW-1 on data share
  // `part`; lane j < CW of the part-0 wave is chain grp*CW + j's lead.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int grp = wave / wpc;
  const int part = wave - grp * wpc;
  const int c_first = grp * CW;
  const int c_local = c_first + (lane < CW ? lane : 0);
  const uint32_t chain = chain0 + (uint32_t)c_local;
  const bool active = chain0 + (uint32_t)c_first < H.n_chains;
  const bool lead = part == 0 && lane < CW && chain < H.n_chains;
  double* st = st_all + c_local * SW;
  float* U = Uall + c_local * NU;
  const float* Uw = Uall + c_first * NU;

  for (int i = tid; i < cpb * NU; i += blockDim.x) Uall[i] = H.u_init[i % NU];
  __syncthreads();
  if (lead) chain_init<typename E::Params>(Ap, chain, st, U);

  const uint32_t tile_len = H.tile_len;
  const uint32_t n4 = tile_len / 4;      // float4s per column of a tile
  const uint32_t tile4 = H.n_cols * n4;  // float4s per tile
  const uint32_t ld4 = H.ld / 4;
  const uint32_t n_tiles = H.n_pad / tile_len;
  const float4* src = reinterpret_cast<const float4*>(H.data);
  const uint32_t wave_base = (uint32_t)(wave * 64);
  float4* res = reinterpret_cast<float4*>(smem + H.lds_tile_offset);

  constexpr int OPT = E::kOpt;
  const TileCopy cp = tile_copy_plan<OPT>(H.n_cols, ld4, n4, tile4, wave_base);
  auto issue_tile = [&](uint32_t ti, float4* dst) {
    issue_tile_copy<E>(src, cp, n4, lane, ti, dst);
  };

  if constexpr (!STREAM) {
    // The whole (padded) data set stays in LDS for every iteration.
    for (uint32_t i = tid; i < tile4; i += blockDim.x)
      res[i] = src[tile_src4<OPT>(i, H.n_cols, ld4, n4, 0)];
  } else {
    if (H.stream_mode == 1 && H.n_vm > 0) issue_tile(0, g_tile_a);
  }
  __syncthreads();

  for (uint32_t it = 0; it <= H.n_iter; ++it) {
    // ---- data pass for the chain's current proposal -------------------
    double acc[CW];
#pragma unroll
    for (int j = 0; j < CW; ++j) acc[j] = 0.0;
    if (H.n_vm > 0) {
      // The uniforms of this pass's proposals, held across its tiles.
      const typename E::AccState ust = E::acc_state(H, Uw);
      if constexpr (!STREAM) {
        if (active)
          tile_pass<E>(H, ust, reinterpret_cast<const float*>(res), 0, part,
                       wpc, Uw, lane, acc);
      } else if (H.stream_mode == 1) {
        // Double buffering over an even number of tiles: tile ti is in A
        // when the pair starts; the copy of ti+1 into B overlaps the pass
        // over A, and the copy of ti+2 (or tile 0 of the next iteration,
        // the data never change) into A overlaps the pass over B.
        for (uint32_t ti = 0; ti < n_tiles; ti += 2) {
          issue_tile(ti + 1, g_tile_b);
          if (active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_a),
                         ti * tile_len, part, wpc, Uw, lane, acc);
          __syncthreads();
          if (ti + 2 < n_tiles)
            issue_tile(ti + 2, g_tile_a);
          else if (it < H.n_iter)
            issue_tile(0, g_tile_a);
          if (active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_b),
                         (ti + 1) * tile_len, part, wpc, Uw, lane, acc);
          __syncthreads();
        }
      } else {
        // Rows too wide for two buffers: synchronous copy into A.
        for (uint32_t ti = 0; ti < n_tiles; ++ti) {
          __syncthreads();
          for (uint32_t i = tid; i < tile4; i += blockDim.x)
            g_tile_a[i] = src[tile_src4<OPT>(i, H.n_cols, ld4, n4, ti)];
          __syncthreads();
          if (active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_a),
                         ti * tile_len, part, wpc, Uw, lane, acc);
        }
      }
    }
    // ---- reduction ----------------------------------------------------
#pragma unroll
    for (int j = 0; j < CW; ++j) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1)
        acc[j] += __shfl_xor(acc[j], off, 64);
      if (lane == 0) partials[(c_first + j) * wpc + part] = acc[j];
    }
    __syncthreads();
    // ---- fused step ---------------------------------------------------
    if (lead) {
      double ll = 0.0;
      for (int p = 0; p < wpc; ++p) ll += partials[c_local * wpc + p];
      if (H.n_vm > 0) ll = E::finish(H, U, ll, true);
      chain_finish_iter(Ap, it, chain, st, ll);
    }
    if (it > 0 && H.swap_every > 0) {
      const uint32_t b = batch_ending_at(it - 1, H.n_iter, H.swap_every,
                                         H.n_batches);
      if (b > 0) {
        // End of a batch: per-chain adaptation, then the replica-exchange
        // sweep of the workgroup's ladder groups, before any proposal of
        // the next batch is drawn.
        const uint32_t bsize =
            b == H.n_batches ? H.n_iter - (H.n_batches - 1) * H.swap_every
                             : H.swap_every;
        __syncthreads();
        if (lead) chain_adapt(Ap, b, bsize, chain, st);
        if (H.swap_group > 1) {
          __syncthreads();
          if (tid == 0) pt_sweep(Ap, b, chain0, cpb, st_all, SW);
        }
        __syncthreads();
      }
    }
    if (part == 0 && active && it < H.n_iter)
      chain_normals(Ap, it, chain0 + (uint32_t)c_first, CW, lane,
                    st_all + c_first * SW, SW);
    __syncthreads();
    if (lead && it < H.n_iter)
      chain_prepare_iter<typename E::Params>(Ap, it, chain, st, U);
    __syncthreads();
  }
  if (lead) chain_finish(Ap, chain, st);
}


// End of batch b (1-based) of one model, rows t0 .. t0 + bsize - 1 of its
// draws, by the whole workgroup (a thread per chain): the batch acceptance
// joins the history; with `adapt` the chain's Welford moments absorb the
// batch's f32 draws and sd = sqrt(max(diag cov, 1e-12)), the seed
// covariance (initial sd^2) until two draws (R/adaptive.R:487-497); the
// population snapshot becomes the last f32 draws, which the next batch
// starts from and draws its DE pairs from (R/adaptive.R:500-502).
__device__ __noinline__ void batch_end_model(const BatchModelInfo* __restrict__ Mp,
                                             uint32_t t0, uint32_t bsize,
                                             uint32_t b, bool adapt) {
  // Separate roundings, as the host orchestrator computes these moments.
#pragma clang fp contract(off)
  const BatchModelInfo M = *Mp;
  const uint32_t np = M.n_params, C = M.n_chains;
  const size_t ld = M.draws_ld;
  for (uint32_t c = threadIdx.x; c < C; c += blockDim.x) {
    if (M.accept_hist) M.accept_hist[(size_t)(b - 1) * C + c] = M.accept[c];
    if (adapt) {
      double* w = M.welford + (size_t)c * (1 + 2 * np);
      double* mean = w + 1;
      double* m2 = w + 1 + np;
      double n = w[0];
      for (uint32_t t = t0; t < t0 + bsize; ++t) {
        n += 1.0;
        for (uint32_t j = 0; j < np; ++j) {
          const double x = (double)M.draws[t + ld * (c + (size_t)C * j)];
          const double d1 = x - mean[j];
          mean[j] += d1 / n;
          m2[j] += d1 * (x - mean[j]);
        }
      }
      w[0] = n;
      for (uint32_t j = 0; j < np; ++j) {
        const double s0 = M.sd0[(size_t)c * np + j];
        const double var = n >= 2.0 ? m2[j] / (n - 1.0) : s0 * s0;
        M.sd[(size_t)c * np + j] = sqrt(fmax(var, 1e-12));
      }
    }
    const uint32_t last = t0 + bsize - 1;
    for (uint32_t j = 0; j < np; ++j)
      M.snapshot[(size_t)c * np + j] =
          (double)M.draws[last + ld * (c + (size_t)C * j)];
  }
}

// Thread 0 waits until *word >= v.  The wait is bounded: past `ticks` of
// the wall clock, or once another workgroup has given up, it raises the
// error word and returns false, so every waiting workgroup ends.
__device__ __noinline__ bool batch_wait_at_least(uint32_t* word, uint32_t v,
                                                 uint32_t* err, uint64_t ticks) {
  if (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= v)
    return true;
  const uint64_t t_start = wall_clock64();
  for (;;) {
    __builtin_amdgcn_s_sleep(8);
    if (__hip_atomic_load(word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= v)
      return true;
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)
      return false;
    if (wall_clock64() - t_start > ticks) {
      __hip_atomic_store(err, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      return false;
    }
  }
}

// Many models, every batch, one persistent launch (BatchQueue).  A
// workgroup takes the next position of its XCD's queue, waits until the
// item's model has ended the previous batch, and runs the item's chain
// group through the batch with the model's arguments (copied to LDS and
// patched with the batch's seed, length and draw offset).  The workgroup
// completing the model's batch runs its between-batch step.
//
// No wait can close a cycle: an item waits only for items of the same
// model's previous batch, which sit earlier in the same queue and were
// taken by workgroups already running (a workgroup takes a position only
// after finishing its previous item), and items of batch 1 wait for
// nothing.  Writes are published with agent-scope release fences before the
// counters move and read after agent-scope acquire fences, so the order
// holds across CUs and XCDs alike.
template <class E>
__device__ __forceinline__ void batch_body(const BatchQueue* __restrict__ Q,
                                           const ChainArgs* __restrict__ args,
                                           const ChainHot* __restrict__ hots) {
  struct Item {
    ChainArgs a;
    uint32_t pos, model, chain0, last;
  };
  __shared__ Item it_sh;
  const uint32_t x = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
  if (slot >= Q->slots[x]) return;  // the whole workgroup leaves at once
  uint32_t* next = Q->sync;
  uint32_t* err = Q->sync + kXcds;
  uint32_t* done = Q->sync + kBatchSyncHead;
  uint32_t* ready = done + Q->n_models;
  const uint32_t len = Q->len[x];
  const uint32_t total = len * Q->n_batches;
  const uint32_t nb = Q->n_batches, every = Q->every;
  const int tid = (int)threadIdx.x;
  for (;;) {
    __syncthreads();  // the previous item is done with it_sh
    if (tid == 0) {
      uint32_t pos = total;
      if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u)
        pos = atomicAdd(&next[x], 1u);
      if (pos < total) {
        const uint32_t b = pos / len + 1;
        const BatchWork w = Q->items[Q->off[x] + (pos - (b - 1) * len)];
        if (!batch_wait_at_least(&ready[w.model], b - 1, err, Q->wait_ticks))
          pos = total;
        it_sh.model = w.model;
        it_sh.chain0 = w.chain0;
      }
      it_sh.pos = pos;
    }
    __syncthreads();
    // Workgroup-uniform values read from LDS: readfirstlane keeps them (and
    // the model's hot values loaded with them) in scalar registers.
    const uint32_t pos = __builtin_amdgcn_readfirstlane(it_sh.pos);
    if (pos >= total) break;  // uniform: every wave leaves here
    // The previous batch's snapshot, scales and Welford moments, written by
    // other workgroups, are read after this fence.
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const uint32_t b = pos / len + 1;
    const uint32_t m = __builtin_amdgcn_readfirstlane(it_sh.model);
    const uint32_t chain0 = __builtin_amdgcn_readfirstlane(it_sh.chain0);
    const uint32_t n_it = b == nb ? Q->n_iter - (nb - 1) * every : every;
    const uint32_t t0 = (b - 1) * every;
    if (tid == 0) {
      it_sh.a = args[m];
      it_sh.a.seed += b - 1;
      it_sh.a.n_iter = n_it;
      it_sh.a.draws += t0;
      it_sh.a.draws_ld = Q->n_iter;
    }
    __syncthreads();
    ChainHot H = hots[m];
    H.n_iter = n_it;
    chain_body<E, true>(&it_sh.a, H, chain0);
    // Publish the group's draws and acceptance, then count the item.
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    __syncthreads();
    if (tid == 0) {
      const uint32_t d = atomicAdd(&done[m], 1u) + 1u;
      it_sh.last = d == b * Q->models[m].n_groups ? 1u : 0u;
    }
    __syncthreads();
    if (__builtin_amdgcn_readfirstlane(it_sh.last)) {
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      batch_end_model(&Q->models[m], t0, n_it, b, b <= Q->warmup_batches);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __syncthreads();
      if (tid == 0)
        __hip_atomic_store(&ready[m], b, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// Data term of a list of parameter points (no chain state).  A persistent
// grid takes work items from a queue: item (g, s) sums observation slice s
// for slots g*cpb .. g*cpb+cpb-1, one wave per slot, and writes one FP64
// partial per (slice, slot); the caller adds the slices in order.  The
// number of slices is chosen on the device from the active slot count so
// that about `target_items` items exist whatever the count, and published
// in *n_slices_out.  The last workgroup to finish resets the queue.
// The uniform table of slot k's point (lead lane of the slot's wave).
template <class E>
__device__ __noinline__ void points_hoist(const PointsArgs* __restrict__ P,
                                          uint32_t k, float* U) {
  const uint32_t pi = P->index ? P->index[k] : k;
  E::hoist(P->hoist, P->n_hoist, P->consts,
           P->points + (size_t)pi * P->n_params, U);
}

// Item count and slicing of one launch (identical in every workgroup).
struct PointsPlan {
  uint32_t count, groups, n_sl, tps, total;
};

__device__ __noinline__ PointsPlan points_plan(const PointsArgs* __restrict__ P,
                                               uint32_t cpb, uint32_t n_tiles) {
  PointsPlan q;
  q.count = P->count ? *P->count : P->n_points;
  q.groups = (q.count + cpb - 1) / cpb;
  // Slices rounded down: no more than target_items items unless the slot
  // groups alone are more, so that a grid of target_items workgroups takes
  // every item in one round (a few items over would run a second round).
  uint32_t n_sl = q.groups ? P->target_items / q.groups : 1u;
  n_sl = max(1u, min(n_sl, n_tiles));
  q.tps = (n_tiles + n_sl - 1) / n_sl;
  q.n_sl = (n_tiles + q.tps - 1) / q.tps;
  q.total = q.groups * q.n_sl;
  return q;
}

template <class E, bool STREAM>
__device__ __forceinline__ void points_body(const PointsArgs* __restrict__ P,
                                            ChainHot H) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ uint32_t item_sh;
  __shared__ PointsPlan plan_sh;
  constexpr int CW = E::kCw, OPT = E::kOpt;
  const int tid = (int)threadIdx.x;
  const int lane = tid & 63;
  // Wave w evaluates slots w*CW .. w*CW + CW-1 of the item's group.
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const uint32_t tile_len = H.tile_len;
  const uint32_t n_tiles = H.n_pad / tile_len;
  if (tid == 0) {
    plan_sh = points_plan(P, H.cpb, n_tiles);
    if (blockIdx.x == 0) *P->n_slices_out = plan_sh.n_sl;
  }
  const int NU = (int)H.n_uniform_pad;
  float* Uall = reinterpret_cast<float*>(smem);
  float* Uw = Uall + wave * CW * NU;
  const uint32_t n4 = tile_len / 4;
  const uint32_t tile4 = H.n_cols * n4;
  const uint32_t ld4 = H.ld / 4;
  const float4* src = reinterpret_cast<const float4*>(H.data);
  const TileCopy cp =
      tile_copy_plan<OPT>(H.n_cols, ld4, n4, tile4, (uint32_t)(wave * 64));
  auto issue_tile = [&](uint32_t ti, float4* dst) {
    issue_tile_copy<E>(src, cp, n4, lane, ti, dst);
  };
  bool loaded = false;  // resident layout: data loaded with the first item
  for (;;) {
    __syncthreads();  // previous item's reads of U and the tiles are done
    if (tid == 0) item_sh = atomicAdd(&P->queue[0], 1u);
    __syncthreads();
    const uint32_t item = item_sh;
    if (item >= plan_sh.total) break;  // uniform: every wave leaves here
    // Slice-major order: the items in flight together read few slices, so
    // the slot groups of one slice share its tiles in the caches.
    const uint32_t groups = plan_sh.groups, tps = plan_sh.tps;
    const uint32_t sl = item / groups, g = item - sl * groups;
    const uint32_t k0 = g * H.cpb + (uint32_t)(wave * CW);
    const bool active = k0 < plan_sh.count;
    const uint32_t kl = k0 + (uint32_t)(lane < CW ? lane : 0);  // lane's slot
    for (int i = tid; i < (int)H.cpb * NU; i += blockDim.x) Uall[i] = H.u_init[i % NU];
    __syncthreads();
    if (lane < CW && kl < plan_sh.count)
      points_hoist<typename E::Params>(P, kl, Uw + lane * NU);
    const uint32_t ti0 = min(sl * tps, n_tiles);
    const uint32_t ti1 = min(ti0 + tps, n_tiles);
    double acc[CW];
#pragma unroll
    for (int j = 0; j < CW; ++j) acc[j] = 0.0;
    if (H.n_vm > 0) {
      // The wave's own lead lanes wrote these uniforms just above.
      const typename E::AccState ust = E::acc_state(H, Uw);
      if constexpr (!STREAM) {
        // One tile holds the whole data set; load it once per workgroup.
        if (!loaded) {
          float4* res = reinterpret_cast<float4*>(smem + H.lds_tile_offset);
          for (uint32_t i = tid; i < tile4; i += blockDim.x)
            res[i] = src[tile_src4<OPT>(i, H.n_cols, ld4, n4, 0)];
          loaded = true;
        }
        __syncthreads();
        if (active && ti0 < ti1)
          tile_pass<E>(H, ust,
                       reinterpret_cast<const float*>(smem + H.lds_tile_offset),
                       0, 0, 1, Uw, lane, acc);
      } else if (H.stream_mode == 1) {
        if (ti0 < ti1) issue_tile(ti0, g_tile_a);
        __syncthreads();
        for (uint32_t ti = ti0; ti < ti1; ti += 2) {
          const bool two = ti + 1 < ti1;
          if (two) issue_tile(ti + 1, g_tile_b);
          if (active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_a),
                         ti * tile_len, 0, 1, Uw, lane, acc);
          __syncthreads();
          if (ti + 2 < ti1) issue_tile(ti + 2, g_tile_a);
          if (two && active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_b),
                         (ti + 1) * tile_len, 0, 1, Uw, lane, acc);
          __syncthreads();
        }
      } else {
        for (uint32_t ti = ti0; ti < ti1; ++ti) {
          __syncthreads();
          for (uint32_t i = tid; i < tile4; i += blockDim.x)
            g_tile_a[i] = src[tile_src4<OPT>(i, H.n_cols, ld4, n4, ti)];
          __syncthreads();
          if (active)
            tile_pass<E>(H, ust, reinterpret_cast<const float*>(g_tile_a),
                         ti * tile_len, 0, 1, Uw, lane, acc);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < CW; ++j) {
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1)
        acc[j] += __shfl_xor(acc[j], off, 64);
      const uint32_t k = k0 + (uint32_t)j;
      // The sum tail's constant joins the first slice's partial only.
      if (lane == 0 && k < plan_sh.count) {
        double v = acc[j];
        if (H.n_vm > 0) v = E::finish(H, Uw + j * NU, v, sl == 0);
        P->partial[(size_t)sl * P->n_points + k] = v;
      }
    }
  }
  // The last workgroup out leaves the queue at zero for the next launch.
  if (tid == 0) {
    __threadfence();
    if (atomicAdd(&P->queue[1], 1u) == gridDim.x - 1) {
      P->queue[0] = 0u;
      P->queue[1] = 0u;
      __threadfence();
    }
  }
}

// The interpreter's parameter programs (one instance for every variant).
struct VmParams {
  static __device__ __forceinline__ void hoist(const uint32_t* code, uint32_t n,
                                               const double* consts,
                                               const double* p, float* U) {
    eval_params64(code, n, consts, p, U);
  }
  static __device__ __forceinline__ double prior(const uint32_t* code,
                                                 uint32_t n,
                                                 const double* consts,
                                                 const double* p) {
    return eval_params64(code, n, consts, p, nullptr);
  }
};

// The interpreter as an evaluator.
template <int NS, int OPT, int CW>
struct VmEval {
  static constexpr int kOpt = OPT;
  static constexpr int kCw = CW;
  static constexpr bool kPairs = NS == 0;
  static constexpr int kUnit = 0;
  static constexpr bool kScalarCopy = false;
  using Params = VmParams;
  // The interpreter reads the uniforms as it goes.
  struct AccState {};
  static __device__ __forceinline__ void run(const ChainHot& H,
                                             const AccState&, const float* U,
                                             const float* tile, uint32_t ld,
                                             int lane, float (&T)[CW][OPT]) {
    vm_run<NS, OPT, CW>((vm_prog_t)H.vm, H.n_vm, U, H.n_uniform_pad, tile, ld,
                        lane, T);
  }
  static __device__ __forceinline__ uint32_t mode(const ChainHot& H) {
    return H.sum_mode;
  }
  static __device__ __forceinline__ AccState acc_state(const ChainHot&,
                                                       const float*) {
    return AccState{};
  }
  static __device__ __forceinline__ void run_acc(const ChainHot& H,
                                                 const AccState&,
                                                 const float* U,
                                                 const float* tile,
                                                 uint32_t ld, int lane,
                                                 float (&s)[CW]) {
    vm_run_acc<OPT, CW>((vm_prog_t)H.vm, H.n_vm, U, H.n_uniform_pad, tile, ld,
                        lane, s);
  }
  static __device__ __forceinline__ double finish(const ChainHot& H,
                                                  const float* U, double s,
                                                  bool with_const) {
    if (H.n_sum_tail == 0) return s;
    double alpha = 1.0, beta = 0.0;
    const VmIns4* tail = H.vm + H.n_vm + 1;
    for (uint32_t i = 0; i < H.n_sum_tail; ++i)
      vm_tail_step(tail[i].op, (double)U[tail[i].a], (double)U[tail[i].b],
                   alpha, beta);
    return vm_tail_apply(alpha, beta, s, H.n_obs, with_const);
  }
};

// The interpreter of affine programs (SUM_AFFINE: a zero load, then one
// FMA_UD per product U[a] * D[b], read from the instruction stream).  It
// carries no stack machine, so several chains per wave keep their few
// registers (the general interpreter holds every chain's stack registers
// through each opcode dispatch and spills), and each LDS read of a column
// feeds kCw chains' FMAs.  Its arithmetic is the general interpreter's on
// such a program: run() starts from U[a0] and applies fmaf(U[a], D[k], T)
// per product, run_acc() is vm_run_acc.
template <int OPT, int CW>
struct VmAffineEval {
  static constexpr int kOpt = OPT;
  static constexpr int kCw = CW;
  static constexpr bool kPairs = true;
  static constexpr int kUnit = 0;
  static constexpr bool kScalarCopy = false;
  using Params = VmParams;
  struct AccState {};
  static __device__ __forceinline__ void run(const ChainHot& H,
                                             const AccState&, const float* U,
                                             const float* tile, uint32_t ld,
                                             int lane, float (&T)[CW][OPT]) {
    const vm_prog_t prog = (vm_prog_t)H.vm;
    const uint32_t nu = H.n_uniform_pad;
    VmIns4 nxt = vm_fetch(prog, 0);
#pragma unroll
    for (int j = 0; j < CW; ++j) {
      const float u0 = U[j * nu + nxt.a];
#pragma unroll
      for (int k = 0; k < OPT; ++k) T[j][k] = u0;
    }
    nxt = vm_fetch(prog, 1);
    for (uint32_t pc = 1; pc < H.n_vm; ++pc) {
      const uint32_t a = nxt.a, b = nxt.b;
      nxt = vm_fetch(prog, pc + 1);
      float D[OPT];
      vm_load_col<OPT>(tile + b * ld, lane, D);
#pragma unroll
      for (int j = 0; j < CW; ++j) {
        const float ua = U[j * nu + a];
#pragma unroll
        for (int k = 0; k < OPT; ++k) T[j][k] = fmaf(ua, D[k], T[j][k]);
      }
    }
  }
  static __device__ __forceinline__ uint32_t mode(const ChainHot&) {
    return SUM_AFFINE;
  }
  static __device__ __forceinline__ AccState acc_state(const ChainHot&,
                                                       const float*) {
    return AccState{};
  }
  static __device__ __forceinline__ void run_acc(const ChainHot& H,
                                                 const AccState&,
                                                 const float* U,
                                                 const float* tile,
                                                 uint32_t ld, int lane,
                                                 float (&s)[CW]) {
    vm_run_acc<OPT, CW>((vm_prog_t)H.vm, H.n_vm, U, H.n_uniform_pad, tile, ld,
                        lane, s);
  }
  static __device__ __forceinline__ double finish(const ChainHot& H,
                                                  const float* U, double s,
                                                  bool with_const) {
    return VmEval<0, OPT, CW>::finish(H, U, s, with_const);
  }
};

}  // namespace gpum
