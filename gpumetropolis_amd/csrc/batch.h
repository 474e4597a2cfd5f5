// batch.h -- host side of a many-model run: the sync words of the
// persistent batched kernel (BatchQueue in chain_kernel.h).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gpum {

// Bound of one wait of the persistent kernel for a model's previous batch.
constexpr uint32_t kBatchWaitSeconds = 30;

// Zero the queue heads, error word and per-model counters before a launch
// (a kernel, not a memset node, so the pair can sit in a HIP graph).
hipError_t launch_batch_sync_reset(uint32_t* sync, uint32_t n_words,
                                   hipStream_t s);

}  // namespace gpum
