// chain_kernel.h -- launch arguments of the fused chain kernel.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vm.h"

namespace gpum {

struct ChainArgs {
  // per-observation VM program (decoded, 4 words per instruction)
  const VmIns4* vm;
  uint32_t n_vm;
  // FP64 hoist program, its constants, and the prior program
  const uint32_t* hoist;
  uint32_t n_hoist;  // words
  const double* consts;
  const uint32_t* prior;
  uint32_t n_prior;  // words
  const double* prior_consts;
  // uniform table template (constants filled, hoisted slots 0), n_uniform_pad
  const float* u_init;
  uint32_t n_uniform_pad;
  // data: FP32 column-major, column stride ld (floats), padded to n_pad
  const float* data;
  uint32_t n_cols;
  uint32_t n_obs;
  uint32_t n_pad;
  uint32_t ld;
  // chains
  uint32_t n_chains;
  uint32_t n_params;
  const double* init;          // chain-major; also the DE snapshot
  const double* proposal_sd;   // chain-major
  const double* temperatures;  // n_chains
  const double* proposal_l;    // n_chains * np * np or null
  double gamma;
  double de_noise;
  uint32_t mode;
  uint32_t n_iter;
  uint32_t seed;
  uint32_t stream_id;
  // outputs
  float* draws;
  float* accept_rate;
  double* last_log_post;
  double* final_state;
  double* init_log_post;
  // Batches inside one launch (parallel tempering and per-chain adaptation,
  // R/adaptive.R:272-413): swap_every = 0 disables both.
  uint32_t iter_offset;    // Philox iteration counter of it = 0
  uint32_t swap_every;
  uint32_t n_batches;
  uint32_t swap_group;     // chains per temperature ladder (<= chains per block)
  uint32_t adapt_batches;  // leading batches with Robbins-Monro adaptation
  uint32_t batch0;         // global index of batch 0 (for resumed runs)
  const double* init_scale;  // per-chain starting RM scale or null (1.0)
  uint8_t* swap_hist;      // n_batches x n_chains: 1/0 per attempted pair
  float* acc_hist;         // n_batches x n_chains: batch acceptance
  double* final_sd;        // n_chains x np
  double* final_scale;     // n_chains
  double* trace_plp;
  double* trace_prop;
  uint8_t* trace_accept;
  double* trace_ll;        // data term of each proposal
  uint32_t draws_ld;       // iteration stride of `draws` (0: n_iter)
  // Non-zero *stop (written by an earlier launch on the stream): the launch
  // does nothing (a batch of a device-side warmup after its early freeze).
  const uint32_t* stop;
  // tiling
  uint32_t cpb;  // chains per workgroup
  uint32_t wpc;  // waves per chain
  uint32_t tile_len;
  uint32_t resident;
  uint32_t lds_tile_offset;
};

// The values the data pass reads every iteration, passed by value so they
// sit in scalar registers; everything else is read through the ChainArgs
// pointer by the lead lane's step code.
struct ChainHot {
  const VmIns4* vm;  // decoded VM program: (op, slot, a, b) per instruction
  const float* u_init;
  const float* data;
  uint32_t n_vm;
  uint32_t n_uniform_pad;
  uint32_t n_cols;
  uint32_t n_obs;
  uint32_t n_pad;
  uint32_t ld;
  uint32_t n_chains;
  uint32_t n_iter;
  uint32_t cpb;  // chains per workgroup (a multiple of the chains per wave)
  uint32_t wpc;  // waves sharing one wave's chains' data pass
  uint32_t tile_len;
  // 0: whole data set resident in LDS; 1: double-buffered streaming through
  // LDS DMA copies; 2: synchronous single-buffer streaming.
  uint32_t stream_mode;
  uint32_t lds_tile_offset;
  uint32_t state_words;
  uint32_t swap_every;
  uint32_t n_batches;
  uint32_t swap_group;
  // Summation of the program (SumMode) and the length of its sum tail,
  // stored in `vm` after the program's padding instruction.
  uint32_t sum_mode;
  uint32_t n_sum_tail;
  // Chunks per FP32 partial (1 or 2): the unit of observations one FP32
  // sum covers before its FP64 add.  Fixed by the model (its columns and
  // observations per lane), so every tiling of a model groups the same
  // observations and rounds them alike.
  uint32_t unit_chunks;
};

// Streamed tiles: two static LDS buffers of kTileF4 float4s each.
constexpr uint32_t kTileF4 = 4608;
constexpr size_t kStreamStaticLds = 2 * (size_t)kTileF4 * 16;
// Copies per wave that fill one buffer with 1024 threads.
constexpr uint32_t kTileIssues = (kTileF4 + 1023) / 1024;

// Batched runs: a work item is one chain group (model, first chain) for one
// batch of iterations.
struct BatchWork {
  uint32_t model;
  uint32_t chain0;
};
// XCDs of an MI355X; the dispatcher hands workgroup w to XCD w mod kXcds.
constexpr uint32_t kXcds = 8;

// Per-model view of the flat per-chain buffers of a batched run.
struct BatchModelInfo {
  uint32_t chain_off;  // first global chain of the model
  uint32_t n_chains;
  uint32_t n_params;
  uint32_t n_groups;     // chain groups (work items) per batch
  uint32_t draws_ld;     // iterations of the whole run
  const float* draws;    // reference layout t + draws_ld (c + C j)
  double* snapshot;      // C x np: start states / DE population
  double* sd;            // C x np proposal scale
  const double* sd0;     // C x np initial proposal scale (Welford seed)
  double* welford;       // C x (1 + 2 np): count, mean, M2
  const float* accept;   // C: acceptance of the batch just run
  float* accept_hist;    // n_batches x C
};

// One persistent launch runs every batch of every model.  Workgroup w
// serves queue x = w mod kXcds (its XCD) as slot w / kXcds; queue x holds
// items[off[x] .. off[x] + len[x]) once per batch, batch-major, so the
// queue position i is item i mod len[x] of batch i / len[x] + 1.  An item
// of batch b starts once batch b-1 of its model has ended (ready[model] >=
// b-1); the workgroup that completes a model's batch runs its between-batch
// step and then publishes ready[model] = b.
// sync: next[kXcds] (queue heads), error word, done[n_models] (items
// completed), ready[n_models] (batches ended); all zero at launch.
struct BatchQueue {
  const BatchWork* items;
  BatchModelInfo* models;
  uint32_t* sync;
  uint32_t off[kXcds];
  uint32_t len[kXcds];
  uint32_t slots[kXcds];    // workgroups serving each queue
  uint32_t n_models;
  uint32_t n_batches;
  uint32_t every;           // iterations per batch; the last absorbs the rest
  uint32_t n_iter;          // iterations of the whole run (draw row stride)
  uint32_t warmup_batches;  // leading batches whose step adapts the scale
  uint32_t pad;
  uint64_t wait_ticks;      // longest wait for a dependency, wall-clock ticks
};
constexpr uint32_t kBatchSyncHead = kXcds + 1;  // words before done[]
// Static LDS of the batched kernel beyond the tile buffers (its copy of
// the model's arguments and the item being run).
constexpr size_t kBatchArgsLds = 512;
// Static LDS of the points kernel beyond the tile buffers (item and plan).
constexpr size_t kPointsExtraLds = 64;
static_assert(sizeof(ChainArgs) + 16 <= kBatchArgsLds, "ChainArgs too large");

// The data term of a list of parameter points, summed per observation
// slice by a persistent grid of n_workgroups that takes (slot group, slice)
// items from a queue; one wave per `chains per wave` slots.
struct PointsArgs {
  const double* points;    // FP64, np per point
  const uint32_t* index;   // slot k evaluates point index[k] (null: point k)
  const uint32_t* count;   // device count of active slots (null: n_points)
  uint32_t n_points;       // slot capacity (partial row stride)
  uint32_t n_params;
  const uint32_t* hoist;   // FP64 hoist program writing the uniform table
  uint32_t n_hoist;
  const double* consts;
  uint32_t target_items;   // work items to aim for (slices * slot groups)
  uint32_t* queue;         // [next item, finished workgroups], zero between launches
  uint32_t* n_slices_out;  // slices of this launch
  double* partial;         // slices x n_points; slices <= min(tiles, target_items)
};

// Host launchers (not seen by the run-time compiler of per-model kernels).
#ifndef __HIPCC_RTC__
// Variant (ns, opt, cw): spill slots, observations per lane, chains per
// wave; a workgroup has h.cpb / cw * h.wpc waves.
hipError_t launch_chain_kernel(const ChainArgs* dev_args, const ChainHot& h,
                               int ns, int opt, int cw, size_t lds_bytes,
                               hipStream_t s);
// `q` is a device copy of the queue; n_wg = kXcds * the most slots.
hipError_t launch_chain_kernel_batch(const BatchQueue* q, uint32_t n_wg,
                                     const ChainArgs* args,
                                     const ChainHot* hots, int ns, int opt,
                                     int cw, size_t lds_bytes, hipStream_t s);
// `p` is a device copy of the arguments.
hipError_t launch_points_kernel(const PointsArgs* p, uint32_t n_workgroups,
                                const ChainHot& h, int ns, int opt, int cw,
                                size_t lds_bytes, hipStream_t s);
hipError_t set_chain_kernel_lds_limit(size_t bytes);
#endif  // __HIPCC_RTC__

}  // namespace gpum
